// Parameter updates of the admixture sweep: UpdateP, UpdateQ (Gibbs and
// Metropolis), alpha, lambda, epsilon, Fst and the data-collection tail.
// Reductions feeding a Metropolis decision are FP64 with a fixed tree
// (block partials in block order, then one block), so results do not depend
// on scheduling.
//
// Every kernel serves a batch of chains: block b of a launch works on chain
// map[b].chain at block coordinates (map[b].bx, map[b].by) of that chain's own
// grid, with the chain's shapes and state addresses taken from cds[chain].
// A single chain is a batch of one.
#pragma once

#include "devmath.h"
#include "sgpu_internal.h"

namespace sgpu {

constexpr int kNT = 256;   // threads per block for the update kernels
// UpdateP sums up to this many allele-count partials (individual tiles) from
// straight-line loads, all in flight at once; more go through a loop.
constexpr int kPDirectParts = 16;

__device__ void lambda_final_block(const ChainDesc &C, const ModelArgs &m, uint32_t rep,
                                   const double *parts, int nparts);
template <int NF>
__device__ void gibbs_q_block(const ChainDesc &C, const ModelArgs &m, int bx, uint32_t rep,
                              int collect);
__device__ void metro_propose_block(const ChainDesc &C, const ModelArgs &m, int bx, uint32_t rep);

// The chain and block coordinates of this block (uniform: scalar loads).
#define SGPU_BLOCK_PROLOGUE()                          \
    const BlockRef br_ = map[blockIdx.x];              \
    const ChainDesc &C = cds[br_.chain];               \
    const ModelArgs m = C.m;                           \
    const int bx = (int)br_.bx;                        \
    (void)bx

// Sum of x[b * stride + off] over b = first, first + step, ... < n by one lane,
// in eight interleaved accumulators (eight loads in flight) added in a fixed tree.
__device__ __forceinline__ double lane_sum8(const SGPU_G double *x, int n, int stride, int off,
                                            int first, int step)
{
    double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int b = first;
    for (; b + 7 * step < n; b += 8 * step) {
#pragma unroll
        for (int u = 0; u < 8; u++) {
            a[u] += x[(size_t)(b + u * step) * stride + off];
        }
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
        if (b + u * step < n) {
            a[u] += x[(size_t)(b + u * step) * stride + off];
        }
    }
    return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

// Sum over b < n of x[b * stride + off] by the 64 lanes of one wave (lane j
// takes b = j, j + 64, ...), then a fixed tree over the lanes; the same order
// whichever wave of a block runs it.  Valid in lane 0.
__device__ __forceinline__ double wave_strided_sum(const SGPU_G double *x, int n, int stride,
                                                   int off)
{
    return wave_sum(lane_sum8(x, n, stride, off, (int)(threadIdx.x & 63), 64));
}

// Column sums of an [n][ncol] row-major table (ncol <= kMaxK + 1) by the 64
// lanes of one wave: lane j takes rows j, j + 64, ... with every column's load
// of a row issued together, then one shuffle tree per column.  out[c] is valid
// in every lane (broadcast from lane 0).
__device__ __forceinline__ void wave_col_sums(const SGPU_G double *x, int n, int ncol,
                                              double *out /* [kMaxK + 1] */)
{
    double a[kMaxK + 1];
#pragma unroll
    for (int c = 0; c <= kMaxK; c++) {
        a[c] = 0.0;
    }
    for (int b = (int)(threadIdx.x & 63); b < n; b += 64) {
        const SGPU_G double *row = x + (size_t)b * ncol;
#pragma unroll
        for (int c = 0; c <= kMaxK; c++) {
            if (c < ncol) {
                a[c] += row[c];
            }
        }
    }
#pragma unroll
    for (int c = 0; c <= kMaxK; c++) {
        out[c] = c < ncol ? __shfl(wave_sum(a[c]), 0, 64) : 0.0;
    }
}

// ---------------------------------------------------------------------------
// Lambda (UpdateLambda.c:8-102).  Partial sums (256 loci per block, loci with
// more than one allele): sum_a log eps (F model), sum_k sum_a log P, or per pop
// sum_a log P (pop-specific); row bx of lam_part.  The decision (one block per
// chain) groups the normaliser terms by allele count (ahist = (value, number
// of loci) pairs).
// ---------------------------------------------------------------------------
__device__ void lambda_partial_block(const ChainDesc &C, const ModelArgs &m, int bx)
{
    const SGPU_G int32_t *nall = g(C.nall);
    const SGPU_G double *P = g(C.P);
    const SGPU_G double *eps = g(C.eps);
    SGPU_G double *part = g(C.lam_part);
    __shared__ double s_red[kNT / 64];
    const int l = bx * kNT + threadIdx.x;
    const int na = (l < m.L) ? nall[l] : 0;
    const int nsums = m.popspecificlambda ? m.K : 1;
    for (int s = 0; s < nsums; s++) {
        double v = 0.0;
        if (na > 1) {
            if (m.popspecificlambda) {
                for (int aa = 0; aa < na; aa++) {
                    v += log_d(P[((size_t)l * m.K + s) * m.A + aa]);
                }
            } else if (m.freqscorr) {
                for (int aa = 0; aa < na; aa++) {
                    v += log_d(eps[(size_t)l * m.A + aa]);
                }
            } else {
                for (int k = 0; k < m.K; k++) {
                    for (int aa = 0; aa < na; aa++) {
                        v += log_d(P[((size_t)l * m.K + k) * m.A + aa]);
                    }
                }
            }
        }
        const double t = block_sum<kNT>(v, s_red);
        if (threadIdx.x == 0) {
            part[(size_t)bx * nsums + s] = t;
        }
    }
}

__global__ __launch_bounds__(kNT) void k_lambda_partial(const ChainDesc *__restrict__ cds,
                                                         const BlockRef *__restrict__ map)
{
    SGPU_BLOCK_PROLOGUE();
    lambda_partial_block(C, m, bx);
}

// UpdateP's allele count of one (locus, pop, allele): the sum of its n_it
// individual-tile partials (plane stride `stride`), read by NP >= n_it
// straight-line loads all in flight at once (loads past the last partial
// repeat it and are not added); the first gamma attempts of the thread's
// Philox counters are formed while the loads are outstanding.
template <int NP>
__device__ __forceinline__ int part_sum_overlapped(const SGPU_G uint16_t *q, size_t stride, int n_it,
                                                   GammaFirst &gf, uint32_t seed, uint32_t chain,
                                                   uint32_t rep, uint32_t l, uint32_t c2)
{
    uint32_t v[NP];
#pragma unroll
    for (int u = 0; u < NP; u++) {
        v[u] = q[(size_t)min(u, n_it - 1) * stride];
    }
    gf = gamma_first(seed, chain, rep, l, c2, PH_P);
    int n = 0;
#pragma unroll
    for (int u = 0; u < NP; u++) {
        n += u < n_it ? (int)v[u] : 0;
    }
    return n;
}

// ---------------------------------------------------------------------------
// UpdateP (UpdateP.c:106-199): one thread per (locus, pop, allele), a block
// holding whole (locus, pop) groups: nall[l] threads each from the chain's work
// list, or A threads each (256 / A groups) on the dense grid.  Dirichlet
// parameters from the allele counts of the last UpdateZ (summed over the
// individual-tile partials, coalesced), one gamma draw per thread, then the
// group's sum in allele order from LDS and LogRDirichletDisc's rule
// (ran.c:1247-1272) including the sum < UNDERFLO uniform fallback.
// Fused: PSum += P (output.c:673-677) when collecting.  In a sweep the launch
// also carries the lambda decision of F-model chains (kRoleLambda: it reads the
// log-eps sums of the last epsilon pass, and neither UpdateP under the F model
// nor UpdateQ reads lambda) and, in Gibbs sweeps, the Gibbs Q update
// (kRoleQGibbs: it reads the counts of the last UpdateZ and alpha, not P).
// Five waves per SIMD, as UpdateP alone needs (93 VGPRs): the lambda decision
// would raise the kernel to 113, and spills only in its own one-block path.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(5))) void k_update_p(
    const ChainDesc *__restrict__ cds, const BlockRef *__restrict__ map,
    const uint32_t *__restrict__ repp, int collect)
{
    SGPU_BLOCK_PROLOGUE();
    const uint32_t rep = *repp;   // sweep index, read on the device (graph replays)
    if (br_.role == kRoleLambda) {
        lambda_final_block(C, m, rep, C.lam_epart, C.e_blocks);
        return;
    }
    if (br_.role == kRoleQGibbs) {
        gibbs_q_block<8>(C, m, bx, rep, collect);   // few loads in flight: UpdateP's registers
        return;
    }
    if (br_.role == kRoleMetroProp) {
        metro_propose_block(C, m, bx, rep);
        return;
    }
    const SGPU_G int32_t *nall = g(C.nall);
    const SGPU_G uint16_t *nApart = g(C.nApart);
    const SGPU_G double *eps = g(C.eps);
    const SGPU_G double *F = g(C.F);
    const SGPU_G double *lambda = g(C.lambda);
    SGPU_G double *P = g(C.P);
    SGPU_G double *PSum = g(C.PSum);
    __shared__ double s_x[kNT];
    const int A = m.A;
    int l, k, aa;
    bool live;
    if (C.pslots) {   // work list: a group's nall[l] threads are adjacent, allele order
        const uint32_t e = g(C.pslots)[(size_t)bx * kNT + threadIdx.x];
        live = e != 0xFFFFFFFFu;
        l = live ? (int)(e >> 13) : 0;
        k = live ? (int)((e >> 8) & 31u) : 0;
        aa = (int)(e & 255u);
    } else {          // dense grid: A threads per (locus, pop) group
        const int gpb = kNT / A;   // groups per block
        const int grp = (int)threadIdx.x / A;
        const int idx = bx * gpb + grp;   // flat (locus, pop)
        aa = (int)threadIdx.x - grp * A;
        live = grp < gpb && idx < m.L * m.K;
        l = live ? idx / m.K : 0;
        k = live ? idx - l * m.K : 0;
    }
    // Every load below is independent of the others and of nall[l]: a thread
    // that is not live reads element 0, and a live one's (l, k, aa) lies inside
    // the L*K*A arrays even when aa >= nall[l]; so the thread issues them all at
    // once and forms the first gamma attempts (Philox block 0 and its
    // Box-Muller pair, which depend on none of them) while they are in flight.
    const int na = live ? nall[l] : 0;
    const bool act = aa < na;
    const size_t pos = ((size_t)l * m.K + k) * A + aa;
    const size_t lpos = live ? pos : 0;
    const size_t stride = (size_t)m.L * m.K * A;
    const int n_it = C.n_it;
    const double Fk = F[live ? k : 0];
    const double pri = m.freqscorr ? eps[live ? (size_t)l * A + aa : 0] : lambda[live ? k : 0];
    const uint32_t c2 = ((uint32_t)k << 16) | (uint32_t)aa;
    const SGPU_G uint16_t *q = nApart + lpos;
    int n = 0;
    GammaFirst gf;
    if (n_it <= 4) {
        n = part_sum_overlapped<4>(q, stride, n_it, gf, m.seed, m.chain, rep, (uint32_t)l, c2);
    } else if (n_it <= 8) {
        n = part_sum_overlapped<8>(q, stride, n_it, gf, m.seed, m.chain, rep, (uint32_t)l, c2);
    } else if (n_it <= kPDirectParts) {
        n = part_sum_overlapped<kPDirectParts>(q, stride, n_it, gf, m.seed, m.chain, rep,
                                               (uint32_t)l, c2);
    } else {
        // eight in flight, exact integer sums
        gf = gamma_first(m.seed, m.chain, rep, (uint32_t)l, c2, PH_P);
        int n0 = 0, n1 = 0, n2 = 0, n3 = 0;
        int it = 0;
        for (; it + 8 <= n_it; it += 8) {
            const SGPU_G uint16_t *qi = q + it * stride;
            n0 += qi[0] + qi[4 * stride];
            n1 += qi[stride] + qi[5 * stride];
            n2 += qi[2 * stride] + qi[6 * stride];
            n3 += qi[3 * stride] + qi[7 * stride];
        }
        for (; it < n_it; it++) {
            n0 += q[it * stride];
        }
        n = (n0 + n1) + (n2 + n3);
    }
    double x = 0.0;
    if (act) {
        const double par = m.freqscorr ? pri * (1.0 - Fk) / Fk + (double)n : pri + (double)n;
        x = gamma_draw_from(par, gf, m.seed, m.chain, rep, (uint32_t)l, c2, PH_P);
    }
    s_x[threadIdx.x] = x;
    __syncthreads();
    if (!act) {
        return;
    }
    double sum = 0.0;
    const int g0 = (int)threadIdx.x - aa;
    for (int j = 0; j < na; j++) {   // allele order, as the host loop adds them
        sum += s_x[g0 + j];
    }
    const double p = sum < kUnderflo ? 1.0 / (double)na : x / sum;
    P[pos] = p;
    if (collect) {
        PSum[pos] += p;
    }
}

// ANCESTDIST (UpdateSums, output.c:691-700): one more count in the box of
// [0,1] that q falls into, NUMBOXES boxes, q = 1 in the last.  Each (i, k) has
// one thread per launch, so the count is a plain read-modify-write.
__device__ __forceinline__ void ancest_box(const ModelArgs &m, SGPU_G int32_t *adist, int i, int k,
                                           double q)
{
    const int nb = m.numboxes;
    int box = (int)(q * (double)nb);
    if (box == nb) {
        box = nb - 1;
    }
    adist[((size_t)i * m.K + k) * nb + box] += 1;
}

// ---------------------------------------------------------------------------
// Dirichlet(alpha + counts) for one individual into q[] (UpdateQAdmixture,
// UpdateQ.c:387-424 / RDirichlet ran.c:348-359), or the Metropolis proposal
// Dirichlet(alpha) (RDirichletDisc, ran.c:1198-1210) when counts is null.
// ---------------------------------------------------------------------------
template <class PA>
__device__ __forceinline__ void dirichlet_q(const ModelArgs &m, PA alpha, const int32_t *cnt,
                                            uint32_t rep, int i,
                                            uint32_t phase, double *q)
{
    double sum = 0.0;
    for (int k = 0; k < m.K; k++) {
        const double par = cnt ? alpha[k] + (double)cnt[k] : alpha[k];
        q[k] = gamma_draw(par, m.seed, m.chain, rep, (uint32_t)i, (uint32_t)k, phase);
        sum += q[k];
    }
    for (int k = 0; k < m.K; k++) {
        q[k] /= sum;
    }
}

// Store Q[i], add QSum, and accumulate the log-Q statistics the alpha update
// reads: per-pop sum of log Q (LogProbQonepop, UpdateAlpha.c:102-111) and the
// total over the clamped Q (LogProbQ, UpdateAlpha.c:157-180).  Block bx
// writes row bx of logq_part; block 0 records the row count.
template <int NT, class PQ>
__device__ __forceinline__ void finish_q(const ModelArgs &m, int bx, int nrows, int i, bool active,
                                         const double *q, PQ Q, PQ QSum, int collect, PQ logq_part,
                                         SGPU_G int32_t *logq_rows, SGPU_G int32_t *adist,
                                         double *s_red)
{
    double tot = 0.0;
    for (int k = 0; k < m.K; k++) {
        if (active) {
            Q[(size_t)i * m.K + k] = q[k];
            if (collect) {
                QSum[(size_t)i * m.K + k] += q[k];
                if (m.ancestdist) {
                    ancest_box(m, adist, i, k, q[k]);
                }
            }
            tot += log_d(q[k] > kSqrtUnder ? q[k] : kSqrtUnder);
        }
    }
    for (int k = 0; k <= m.K; k++) {
        double v = 0.0;
        if (active) {
            v = (k < m.K) ? log_d(q[k]) : tot;
        }
        const double s = block_sum<NT>(v, s_red);
        if (threadIdx.x == 0) {
            logq_part[(size_t)bx * (m.K + 1) + k] = s;
        }
    }
    if (bx == 0 && threadIdx.x == 0) {
        logq_rows[0] = nrows;
    }
}

constexpr int kRowSplit = 16;      // second-level partials of the ancestry counts

// First level of the ancestry-count reduction: the n_lt locus-tile partials of
// UpdateZ ([n_lt][N*K]) are folded into kRowSplit partials; chain grid (flat
// chunks, kRowSplit), each thread sums every kRowSplit-th row of one flat
// index with four independent accumulators.
__device__ __forceinline__ void fold_rows_block(const ChainDesc &C, const ModelArgs &m, int bx,
                                                int by)
{
    const SGPU_G uint16_t *part = g(C.nQpart);
    SGPU_G int32_t *out = g(C.nQfold);
    const int nrows = C.n_lt;
    const size_t n = (size_t)m.N * m.K;
    const size_t f = (size_t)bx * kNT + threadIdx.x;
    if (f >= n) {
        return;
    }
    int32_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int t = by;
    for (; t + 3 * kRowSplit < nrows; t += 4 * kRowSplit) {
        a0 += part[(size_t)t * n + f];
        a1 += part[(size_t)(t + kRowSplit) * n + f];
        a2 += part[(size_t)(t + 2 * kRowSplit) * n + f];
        a3 += part[(size_t)(t + 3 * kRowSplit) * n + f];
    }
    for (; t < nrows; t += kRowSplit) {
        a0 += part[(size_t)t * n + f];
    }
    out[(size_t)by * n + f] = (a0 + a1) + (a2 + a3);
}

__global__ __launch_bounds__(kNT) void k_fold_rows(const ChainDesc *__restrict__ cds,
                                                   const BlockRef *__restrict__ map)
{
    SGPU_BLOCK_PROLOGUE();
    fold_rows_block(C, m, bx, (int)br_.by);
}

// Gibbs update of Q (UpdateQAdmixture, UpdateQ.c:387-424; RDirichlet,
// ran.c:348-359): one thread per (individual, pop), a block holding whole
// individuals (K * (256 / K) threads).  Each thread sums its folded ancestry
// counts ([kRowSplit][N][K], loads in flight) and draws one gamma; the
// individual's sum is taken in pop order from LDS.  Fused: QSum, the ANCESTDIST
// histogram, and the log-Q statistics of the alpha update, one row per block: per pop the sum
// over the block's individuals of log Q[i][k] (LogProbQonepop,
// UpdateAlpha.c:102-111), and the sum of the per-individual totals of
// log max(Q, 1e-50) (LogProbQ, UpdateAlpha.c:157-180).
template <int NF>
__device__ void gibbs_q_block(const ChainDesc &C, const ModelArgs &m, int bx, uint32_t rep,
                              int collect)
{
    static_assert(NF % 4 == 0 && kQDirectRows % NF == 0, "partials in flight");
    const SGPU_G double *alpha = g(C.alpha);
    SGPU_G double *Q = g(C.Q);
    SGPU_G double *QSum = g(C.QSum);
    SGPU_G double *logq_part = g(C.logq_part);
    __shared__ double s_x[kNT], s_l[kNT], s_c[kNT];
    const int K = m.K;
    const int ipb = kNT / K;   // individuals per block
    const int j = (int)threadIdx.x / K, k = (int)threadIdx.x - j * K;
    const int i = bx * ipb + j;
    const bool active = j < ipb && i < m.N;
    double x = 0.0;
    if (active) {
        const size_t n = (size_t)m.N * K;
        int32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
        if (C.q_direct) {   // few locus tiles (<= kQDirectRows): NF partials in flight
            const SGPU_G uint16_t *f = g(C.nQpart) + (size_t)i * K + k;
            const int nrows = C.n_lt;
#pragma unroll 1
            for (int t0 = 0; t0 < nrows; t0 += NF) {
                int32_t c[NF];
#pragma unroll
                for (int t = 0; t < NF; t++) {
                    c[t] = t0 + t < nrows ? (int32_t)f[(size_t)(t0 + t) * n] : 0;
                }
#pragma unroll
                for (int t = 0; t < NF; t += 4) {
                    c0 += c[t];
                    c1 += c[t + 1];
                    c2 += c[t + 2];
                    c3 += c[t + 3];
                }
            }
        } else {            // the kRowSplit rows of the fold
            const SGPU_G int32_t *f = g(C.nQfold) + (size_t)i * K + k;
#pragma unroll
            for (int s2 = 0; s2 < kRowSplit; s2 += 4) {
                c0 += f[(size_t)s2 * n];
                c1 += f[(size_t)(s2 + 1) * n];
                c2 += f[(size_t)(s2 + 2) * n];
                c3 += f[(size_t)(s2 + 3) * n];
            }
        }
        x = gamma_draw(alpha[k] + (double)((c0 + c1) + (c2 + c3)), m.seed, m.chain, rep,
                       (uint32_t)i, (uint32_t)k, PH_Q);
    }
    s_x[threadIdx.x] = x;
    __syncthreads();
    double q = 0.0;
    if (active) {
        double sum = 0.0;
        for (int kk = 0; kk < K; kk++) {   // pop order, as RDirichlet adds them
            sum += s_x[threadIdx.x - k + kk];
        }
        q = x / sum;
        Q[(size_t)i * K + k] = q;
        if (collect) {
            QSum[(size_t)i * K + k] += q;
            if (m.ancestdist) {
                ancest_box(m, g(C.adist), i, k, q);
            }
        }
    }
    s_l[threadIdx.x] = active ? log_d(q) : 0.0;
    s_c[threadIdx.x] = active ? log_d(q > kSqrtUnder ? q : kSqrtUnder) : 0.0;
    __syncthreads();
    // per individual, the clamped total in pop order (into s_x, now free) ...
    if (k == 0 && j < ipb) {
        double tot = 0.0;
        for (int kk = 0; kk < K; kk++) {
            tot += s_c[threadIdx.x + kk];
        }
        s_x[j] = tot;
    }
    __syncthreads();
    // ... then a pairwise tree over the block's individuals, per pop and for the totals
    for (int st = 1; st < ipb; st *= 2) {
        if (j % (2 * st) == 0 && j + st < ipb) {
            s_l[threadIdx.x] += s_l[threadIdx.x + st * K];
            if (k == 0) {
                s_x[j] += s_x[j + st];
            }
        }
        __syncthreads();
    }
    if ((int)threadIdx.x <= K) {
        logq_part[(size_t)bx * (K + 1) + threadIdx.x] =
            (int)threadIdx.x < K ? s_l[threadIdx.x] : s_x[0];
    }
    if (bx == 0 && threadIdx.x == 0) {
        g(C.logq_rows)[0] = (m.N + ipb - 1) / ipb;
    }
}

__global__ __launch_bounds__(kNT) void k_update_q_gibbs(const ChainDesc *__restrict__ cds,
                                                         const BlockRef *__restrict__ map,
                                                         const uint32_t *__restrict__ repp,
                                                         int collect)
{
    SGPU_BLOCK_PROLOGUE();
    gibbs_q_block<kQDirectRows>(C, m, bx, *repp, collect);
}

// Log-Q statistics of the current Q (after initialisation or a host write of Q).
__global__ __launch_bounds__(kNT) void k_logq_stats(const ChainDesc *__restrict__ cds,
                                                    const BlockRef *__restrict__ map)
{
    SGPU_BLOCK_PROLOGUE();
    const SGPU_G double *Q = g(C.Q);
    SGPU_G double *logq_part = g(C.logq_part);
    __shared__ double s_red[kNT / 64];
    const int i = bx * kNT + threadIdx.x;
    const bool active = i < m.N;
    double q[kMaxK];
    for (int k = 0; k < m.K; k++) {
        q[k] = active ? Q[(size_t)i * m.K + k] : 1.0;
    }
    double tot = 0.0;
    for (int k = 0; k < m.K; k++) {
        tot += active ? log_d(q[k] > kSqrtUnder ? q[k] : kSqrtUnder) : 0.0;
    }
    for (int k = 0; k <= m.K; k++) {
        const double v = active ? ((k < m.K) ? log_d(q[k]) : tot) : 0.0;
        const double s = block_sum<kNT>(v, s_red);
        if (threadIdx.x == 0) {
            logq_part[(size_t)bx * (m.K + 1) + k] = s;
        }
    }
    if (bx == 0 && threadIdx.x == 0) {
        g(C.logq_rows)[0] = C.q_blocks;
    }
}

// ---------------------------------------------------------------------------
// UpdateQMetro (UpdateQ.c:199-307): proposal from the prior, accepted with
// probability exp(sum over copies of log(TestQ.P) - log(Q.P))
// (CalcLikeIndDiff, CalcLike.c:281-345).  Pass 0 draws each individual's
// proposal once into qprop; pass 1 (one thread per individual and a locus
// range per grid row) accumulates both products in FP64 with exponent
// bookkeeping; pass 2 sums the ranges and decides.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kNT) void k_metro_propose(const ChainDesc *__restrict__ cds,
                                                        const BlockRef *__restrict__ map,
                                                        const uint32_t *__restrict__ repp)
{
    SGPU_BLOCK_PROLOGUE();
    const uint32_t rep = *repp;   // sweep index, read on the device (graph replays)
    if (br_.role == kRoleLamPart) {  // lambda log sums of the new P (no F model)
        lambda_partial_block(C, m, bx);
        return;
    }
    metro_propose_block(C, m, bx, rep);
}

// One block of the Metropolis proposals TestQ ~ Dirichlet(alpha) (UpdateQMetro,
// UpdateQ.c:199-307): whole individuals, one thread per (individual, pop).  They
// read only alpha, so in a sweep they run as blocks of the UpdateP launch.
__device__ void metro_propose_block(const ChainDesc &C, const ModelArgs &m, int bx, uint32_t rep)
{
    __shared__ double s_x[kNT];
    const int K = m.K;
    const int ipb = kNT / K;   // one thread per (individual, pop), whole individuals per block
    const int j = (int)threadIdx.x / K, k = (int)threadIdx.x - j * K;
    const int i = bx * ipb + j;
    const bool active = j < ipb && i < m.N;
    double x = 0.0;
    if (active) {
        x = gamma_draw(g(C.alpha)[k], m.seed, m.chain, rep, (uint32_t)i, (uint32_t)k, PH_QMETRO);
    }
    s_x[threadIdx.x] = x;
    __syncthreads();
    if (active) {
        double sum = 0.0;
        for (int kk = 0; kk < K; kk++) {   // pop order, as RDirichletDisc adds them
            sum += s_x[threadIdx.x - k + kk];
        }
        g(C.qprop)[(size_t)i * K + k] = x / sum;
    }
}

// Loci of P staged in LDS at a time by the 2-bit Metropolis ratio: 256 up to
// K = 16 (64 KiB, the most one block may hold), 128 above.
template <int K>
constexpr int metro_loci() { return K <= 16 ? 256 : 128; }

// Biallelic 2-bit path, K templated: the block's P columns are staged in LDS
// allele-major ([locus][allele][pop]), four loci per genotype load; each copy
// reads its own allele's K frequencies (the lanes of a wave read at most two
// addresses, in different banks) and forms its two dot products with fused
// multiply-adds (the ratio feeds an exp() compared with a uniform; it needs no
// reference-exact rounding).
template <int K>
__global__ __launch_bounds__(kNT) void k_metro_ratio_snp2(const ChainDesc *__restrict__ cds,
                                                           const BlockRef *__restrict__ map,
                                                           const uint32_t *__restrict__ /*repp*/)
{
    SGPU_BLOCK_PROLOGUE();
    const int by = (int)br_.by;
    const SGPU_G uint32_t *geno = g(C.geno);
    const SGPU_G double *Q = g(C.Q);
    const SGPU_G double *P = g(C.P);
    const SGPU_G double *qprop = g(C.qprop);
    SGPU_G double *delta_part = g(C.delta_part);
    const int Lpad = C.Lpad, loci_per_range = C.loci_per_range;
    const int i = bx * kNT + threadIdx.x;
    const bool active = i < m.N;
    double tq[K], cq[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        tq[k] = active ? qprop[(size_t)i * K + k] : 1.0;
        cq[k] = active ? Q[(size_t)i * K + k] : 1.0;
    }
    constexpr int kMetroLoci = metro_loci<K>();
    constexpr int KP = (K + 1) & ~1;   // an allele's row padded to 16 bytes (paired reads)
    __shared__ __align__(16) double s_p[kMetroLoci * KP * 2];
    const int l0 = by * loci_per_range;
    const int l1 = min(l0 + loci_per_range, m.L);
    const SGPU_G uint32_t *row = geno + (size_t)(active ? (i >> 4) : 0) * Lpad;
    const int sh = 2 * (i & 15);
    double pp = 1.0, pm = 1.0;
    int ep = 0, em = 0;
    for (int c0 = l0; c0 < l1; c0 += kMetroLoci) {
        const int c1 = min(c0 + kMetroLoci, l1);
        __syncthreads();
        for (int t = threadIdx.x; t < (c1 - c0) * K * 2; t += kNT) {
            // P[l][k][a] -> s_p[l - c0][a][k], rows of KP
            const int lt = t / (2 * K), r = t - lt * 2 * K;
            s_p[lt * 2 * KP + (r & 1) * KP + (r >> 1)] = P[(size_t)c0 * K * 2 + t];
        }
        __syncthreads();
        // four loci per load; the next four are loaded one iteration ahead
        uint4 nw4 = *(const SGPU_G uint4 *)(row + c0);
        for (int lb = c0; lb < c1; lb += 4) {
            const uint4 w4 = nw4;
            if (lb + 4 < c1) {
                nw4 = *(const SGPU_G uint4 *)(row + lb + 4);
            }
            const uint32_t ws[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int l = lb + u;
                const uint32_t code = (ws[u] >> sh) & 3u;
                if (l < c1 && code != 3u) {
                    // copy 0 carries allele 1 only in code 2, copy 1 in codes 1 and 2
                    const double *Pl = s_p + (l - c0) * KP * 2;
                    const double *p0 = Pl + (code == 2u ? KP : 0);
                    const double *p1 = Pl + (code != 0u ? KP : 0);
                    double t0 = 0.0, t1 = 0.0, m0 = 0.0, m1 = 0.0;
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const double a = p0[k], b = p1[k];
                        t0 = fma(tq[k], a, t0);
                        t1 = fma(tq[k], b, t1);
                        m0 = fma(cq[k], a, m0);
                        m1 = fma(cq[k], b, m1);
                    }
                    pp *= t0;
                    pp *= t1;
                    pm *= m0;
                    pm *= m1;
                }
            }
            int e;
            pp = frexp(pp, &e);
            ep += e;
            pm = frexp(pm, &e);
            em += e;
        }
    }
    if (active) {
        delta_part[(size_t)by * m.N + i] = (log_d(pp) - log_d(pm)) + (double)(ep - em) * kLn2;
    }
}

// Byte-per-copy path (any allele count), K templated.  The frequencies of a
// chunk of loci are staged in LDS allele-major ([locus][allele][pop]): a copy
// reads its allele's K values at constant offsets, and the block reads each
// frequency from global memory once instead of once per lane.
// Doubles of staged frequencies: 32 KiB, or one locus of MAXALLELES x K where
// that is more (K > 16 at 255 alleles: up to 40 KiB).
template <int K>
constexpr int metro_byte_lds() { return kMaxAlleles * K > 4096 ? kMaxAlleles * K : 4096; }

template <int K>
__global__ __launch_bounds__(kNT) void k_metro_ratio_byte(const ChainDesc *__restrict__ cds,
                                                           const BlockRef *__restrict__ map,
                                                           const uint32_t *__restrict__ /*repp*/)
{
    SGPU_BLOCK_PROLOGUE();
    const int by = (int)br_.by;
    const SGPU_G uint32_t *geno = g(C.geno);
    const SGPU_G double *Q = g(C.Q);
    const SGPU_G double *P = g(C.P);
    const SGPU_G double *qprop = g(C.qprop);
    SGPU_G double *delta_part = g(C.delta_part);
    const int Gpad = C.Gpad, loci_per_range = C.loci_per_range, A = m.A;
    const int i = bx * kNT + threadIdx.x;
    const bool active = i < m.N;
    double tq[K], cq[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        tq[k] = active ? qprop[(size_t)i * K + k] : 1.0;
        cq[k] = active ? Q[(size_t)i * K + k] : 1.0;
    }
    constexpr int kMetroByteLds = metro_byte_lds<K>();
    __shared__ double s_p[kMetroByteLds];
    const int chunk = kMetroByteLds / (A * K);   // >= 1 (A <= kMaxAlleles)
    const int l0 = by * loci_per_range;
    const int l1 = min(l0 + loci_per_range, m.L);
    const int sh = 16 * (i & 1);
    double pp = 1.0, pm = 1.0;
    int ep = 0, em = 0;
    for (int c0 = l0; c0 < l1; c0 += chunk) {
        const int nc = min(chunk, l1 - c0);
        __syncthreads();
        // P[l][k][a] -> s_p[l - c0][a][k], one (locus, allele) row of K per thread
        for (int r = threadIdx.x; r < nc * A; r += kNT) {
            const int j = r / A, a = r - j * A;
            const SGPU_G double *src = P + (size_t)(c0 + j) * K * A + a;
#pragma unroll
            for (int k = 0; k < K; k++) {
                s_p[r * K + k] = src[k * A];
            }
        }
        __syncthreads();
        if (!active) {
            continue;
        }
        // words [L][Gpad]: lanes 2j and 2j+1 read the same word, adjacent lanes
        // adjacent words; the next locus's word is loaded one iteration ahead
        uint32_t next = geno[(size_t)c0 * Gpad + (i >> 1)];
        for (int l = c0; l < c0 + nc; l++) {
            const uint32_t word = next;
            if (l + 1 < c0 + nc) {
                next = geno[(size_t)(l + 1) * Gpad + (i >> 1)];
            }
            const int xa = (int)((word >> sh) & 0xffu);
            const int xb = (int)((word >> (sh + 8)) & 0xffu);
            const double *pl = s_p + (size_t)(l - c0) * A * K;
            if (xa != 0xff) {
                const double *px = pl + xa * K;
                double tp = 0.0, tm = 0.0;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    tp = fma(tq[k], px[k], tp);
                    tm = fma(cq[k], px[k], tm);
                }
                pp *= tp;
                pm *= tm;
            }
            if (xb != 0xff) {
                const double *px = pl + xb * K;
                double tp = 0.0, tm = 0.0;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    tp = fma(tq[k], px[k], tp);
                    tm = fma(cq[k], px[k], tm);
                }
                pp *= tp;
                pm *= tm;
            }
            int e;
            pp = frexp(pp, &e);
            ep += e;
            pm = frexp(pm, &e);
            em += e;
        }
    }
    if (active) {
        delta_part[(size_t)by * m.N + i] = (log_d(pp) - log_d(pm)) + (double)(ep - em) * kLn2;
    }
}

// Individuals per Metropolis acceptance block: each takes four waves, wave h
// summing the ranges r = h, h + 4, ... of its individuals' log-ratio partials.
constexpr int kAcceptInds = 64;

__global__ __launch_bounds__(kNT) void k_metro_accept(const ChainDesc *__restrict__ cds,
                                                       const BlockRef *__restrict__ map,
                                                       const uint32_t *__restrict__ repp,
                                                       int collect)
{
    SGPU_BLOCK_PROLOGUE();
    const uint32_t rep = *repp;   // sweep index, read on the device (graph replays)
    const SGPU_G double *delta_part = g(C.delta_part);
    SGPU_G double *Q = g(C.Q);
    __shared__ double s_red[kNT / 64];
    __shared__ double s_d[kNT];
    const int li = (int)(threadIdx.x & 63), h = (int)(threadIdx.x >> 6);
    const int i = bx * kAcceptInds + li;
    const bool inds = i < m.N;
    // a quarter of the range partials per wave, sixteen loads in flight per lane
    double a[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
        a[u] = 0.0;
    }
    if (inds) {
        int r = h;
        for (; r + 15 * 4 < C.n_ranges; r += 16 * 4) {
#pragma unroll
            for (int u = 0; u < 16; u++) {
                a[u] += delta_part[(size_t)(r + 4 * u) * m.N + i];
            }
        }
#pragma unroll
        for (int u = 0; u < 16; u++) {
            if (r + 4 * u < C.n_ranges) {
                a[u] += delta_part[(size_t)(r + 4 * u) * m.N + i];
            }
        }
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
        a[u] += a[u + 8];
    }
    s_d[threadIdx.x] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    __syncthreads();
    // wave 0: one individual per lane, the quarters added in wave order
    const bool active = h == 0 && inds;
    double q[kMaxK];
    if (active) {
        const double delta = (s_d[li] + s_d[64 + li]) + (s_d[128 + li] + s_d[192 + li]);
        const u32x4 w = philox4x32_10(rep, (uint32_t)i, 0xffffffffu, PH_QMETRO << 24, m.seed,
                                      m.chain);
        const bool acc = uo(w.x) < exp(delta);
        const SGPU_G double *src = acc ? g(C.qprop) : Q;
        for (int k = 0; k < m.K; k++) {
            q[k] = src[(size_t)i * m.K + k];
        }
    }
    finish_q<kNT>(m, bx, (m.N + kAcceptInds - 1) / kAcceptInds, i, active, q, Q, g(C.QSum),
                  collect, g(C.logq_part), g(C.logq_rows), g(C.adist), s_red);
}

// ---------------------------------------------------------------------------
// UpdateAlpha (UpdateAlpha.c:189-253), one wave per chain: sums the log-Q
// statistics and runs the pop loop in lane 0.  LogProbQ(a) = (a-1) * sum
// log(clamped Q) + (lgamma(K a) - K lgamma(a)) * N; LogProbQonepop as in the
// reference.  As there, logprobdiff is zeroed once and only reset by the
// gamma prior.
// ---------------------------------------------------------------------------
__device__ void alpha_wave(const ChainDesc &C, const ModelArgs &m, uint32_t rep)
{
    const SGPU_G double *logq_part = g(C.logq_part);
    SGPU_G double *alpha = g(C.alpha);
    const int nparts = g(C.logq_rows)[0];
    __shared__ double s_stat[kMaxK + 1];   // written and read by lane 0 only
    {
        double sums[kMaxK + 1];
        wave_col_sums(logq_part, nparts, m.K + 1, sums);
        if ((threadIdx.x & 63) != 0) {
            return;
        }
#pragma unroll
        for (int c = 0; c <= kMaxK; c++) {
            s_stat[c] = sums[c];
        }
    }
    const int K = m.K;
    const double N = (double)m.N;
    const int numalphas = m.popalphas ? K : 1;
    double logprobdiff = 0.0;
    for (int pop = 0; pop < numalphas; pop++) {
        const u32x4 w = philox4x32_10(rep, (uint32_t)pop, 0u, PH_ALPHA << 24, m.seed, m.chain);
        const double z = box_muller(w.x, w.y);
        const double oldalpha = alpha[pop];
        const double newalpha = oldalpha + m.alphapropsd * z;
        if ((newalpha > 0) && ((newalpha < m.alphamax) || (!m.unifprioralpha))) {
            if (!m.unifprioralpha) {
                logprobdiff = (m.alphapriora - 1) * log_d(newalpha / oldalpha) +
                              (oldalpha - newalpha) / m.alphapriorb;
            }
            if (m.popalphas) {
                double sumalphas = 0.0;
                for (int k = 0; k < K; k++) {
                    sumalphas += alpha[k];
                }
                const double t = s_stat[pop];
                logprobdiff -= (oldalpha - 1.0) * t +
                               (lgamma_lanczos(sumalphas) - lgamma_lanczos(oldalpha)) * N;
                sumalphas += newalpha - oldalpha;
                logprobdiff += (newalpha - 1.0) * t +
                               (lgamma_lanczos(sumalphas) - lgamma_lanczos(newalpha)) * N;
            } else {
                const double t = s_stat[K];
                logprobdiff += (newalpha - 1.0) * t +
                               (lgamma_lanczos(K * newalpha) - K * lgamma_lanczos(newalpha)) * N;
                logprobdiff -= (oldalpha - 1.0) * t +
                               (lgamma_lanczos(K * oldalpha) - K * lgamma_lanczos(oldalpha)) * N;
            }
            if (uo(w.z) < exp(logprobdiff)) {
                alpha[pop] = newalpha;
                if (!m.popalphas) {
                    for (int k = 1; k < K; k++) {
                        alpha[k] = newalpha;
                    }
                }
            }
        }
    }
}

__global__ __launch_bounds__(kNT) void k_update_alpha(const ChainDesc *__restrict__ cds,
                                                       const BlockRef *__restrict__ map,
                                                       const uint32_t *__restrict__ repp)
{
    SGPU_BLOCK_PROLOGUE();
    if (threadIdx.x < 64) {
        alpha_wave(C, m, *repp);
    }
}

// The lambda decision of one chain by a whole block of kNT threads: sums the
// nparts rows of partials (eight loads in flight per thread), then per proposal
// the normaliser terms of the allele-count groups in parallel (thread h takes
// groups h, h + kNT, ...) and thread 0 adds them in thread order and decides.
__device__ void lambda_final_block(const ChainDesc &C, const ModelArgs &m, uint32_t rep,
                                   const double *parts, int nparts)
{
    const SGPU_G double *part = g(parts);
    const SGPU_G int32_t *ahist = g(C.ahist);
    SGPU_G double *lambda = g(C.lambda);
    const int nhist = C.nhist;
    __shared__ double s_red[kNT / 64];
    __shared__ double s_stat[kMaxK];
    __shared__ double s_t[kNT];
    const int nsums = m.popspecificlambda ? m.K : 1;
    for (int s = 0; s < nsums; s++) {
        const double t = block_sum<kNT>(
            lane_sum8(part, nparts, nsums, s, (int)threadIdx.x, kNT), s_red);
        if (threadIdx.x == 0) {
            s_stat[s] = t;
        }
    }
    const double stoppop = m.popspecificlambda ? 1.0 : (m.freqscorr ? 1.0 : (double)m.K);
    for (int s = 0; s < nsums; s++) {
        const u32x4 w = philox4x32_10(rep, (uint32_t)s, 0u, PH_LAMBDA << 24, m.seed, m.chain);
        const double z = box_muller(w.x, w.y);
        const double old = lambda[s];
        const double nw = old + kLambdaPropSd * z;
        if (!((nw > 0.0) && (nw < kLambdaMax))) {   // block-uniform
            continue;
        }
        // one Lanczos sum at a time (the rolled loops), which keeps this block
        // within the register budget of the launches that carry it
        double loc = 0.0;
        if ((int)threadIdx.x < nhist) {
            double lg_old = 0.0, lg_nw = 0.0;   // the terms that do not depend on the group
#pragma unroll 1
            for (int w = 0; w < 2; w++) {
                const double v = lgamma_lanczos(w == 0 ? old : nw);
                lg_old = w == 0 ? v : lg_old;
                lg_nw = w == 0 ? lg_nw : v;
            }
#pragma unroll 1
            for (int h = (int)threadIdx.x; h < nhist; h += kNT) {
                const double na = (double)ahist[2 * h];
                const double cnt = (double)ahist[2 * h + 1];
                double t = 0.0;
#pragma unroll 1
                for (int w = 0; w < 2; w++) {
                    const double v = stoppop * lgamma_lanczos(na * (w == 0 ? nw : old));
                    t = w == 0 ? t + v : t - v;
                }
                t += stoppop * na * lg_old;
                t -= stoppop * na * lg_nw;
                loc += cnt * t;
            }
        }
        s_t[threadIdx.x] = loc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double sum = 0.0;
            for (int h = 0; h < nhist && h < kNT; h++) {
                sum += s_t[h];
            }
            sum += (nw - old) * s_stat[s];
            if (uo(w.z) < exp(sum)) {
                if (m.popspecificlambda) {
                    lambda[s] = nw;
                } else {
                    for (int k = 0; k < m.K; k++) {
                        lambda[k] = nw;
                    }
                }
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kNT) void k_lambda_finalize(const ChainDesc *__restrict__ cds,
                                                          const BlockRef *__restrict__ map,
                                                          const uint32_t *__restrict__ repp)
{
    SGPU_BLOCK_PROLOGUE();
    lambda_final_block(C, m, *repp, C.lam_part, C.l_blocks);
}

// Fst proposal of group pop1 (UpdateFst.c:160-166): F + FPRIORSD * N(0,1).
__device__ __forceinline__ double fst_proposal(const ModelArgs &m, double oldf, int pop1,
                                               uint32_t rep, uint32_t *uw)
{
    const u32x4 w = philox4x32_10(rep, (uint32_t)pop1, 0u, PH_FST << 24, m.seed, m.chain);
    if (uw) {
        *uw = w.z;
    }
    return oldf + m.fpriorsd * box_muller(w.x, w.y);
}

// Column sums of an [n][ncol] row-major table (ncol <= kMaxK + 1) by a block of
// NT threads.  With ncol <= NT / 64 each column gets NT / 64 / ncol waves, each
// wave a strided share of the rows reduced by its shuffle tree, and a column's
// wave sums are added in wave order; otherwise wave w takes columns w, w + NT / 64,
// ... whole.  The order is fixed, so the sums do not depend on scheduling.
// out[c] (LDS) is valid in every thread on return.
template <int NT>
__device__ void block_col_sums(const SGPU_G double *x, int n, int ncol, double *s_part /* NT/64 */,
                               double *out /* kMaxK + 1 */)
{
    constexpr int NW = NT / 64;
    const int lane = (int)(threadIdx.x & 63), wid = (int)(threadIdx.x >> 6);
    if (ncol <= NW) {
        const int wpc = NW / ncol;   // waves per column
        const int c = wid / wpc, share = wid - c * wpc;
        double v = 0.0;
        if (c < ncol) {
            v = wave_sum(lane_sum8(x, n, ncol, c, share * 64 + lane, wpc * 64));
        }
        if (lane == 0) {
            s_part[wid] = v;
        }
        __syncthreads();
        if ((int)threadIdx.x < ncol) {
            double t = 0.0;
            for (int s = 0; s < wpc; s++) {
                t += s_part[threadIdx.x * wpc + s];
            }
            out[threadIdx.x] = t;
        }
    } else {
        for (int c = wid; c < ncol; c += NW) {
            const double v = wave_sum(lane_sum8(x, n, ncol, c, lane, 64));
            if (lane == 0) {
                out[c] = v;
            }
        }
    }
    __syncthreads();
}

// lnP(D) of the last UpdateZ as the sum of its tile partials: share s of ns
// contiguous shares, by one wave (valid in lane 0).
__device__ __forceinline__ double like_share(const ChainDesc &C, int s, int ns)
{
    const int n = C.n_it * C.n_lt;
    const int b0 = (int)((long long)n * s / ns), b1 = (int)((long long)n * (s + 1) / ns);
    return wave_sum(lane_sum8(g(C.likepart) + b0, b1 - b0, 1, 0, (int)(threadIdx.x & 63), 64));
}

// Flags of the epsilon launch (kEpsAdvance: it closes the sweep, see k_update_eps).
constexpr int kEpsCollect = 1, kEpsFst = 2, kEpsAlpha = 4, kEpsLike = 8, kEpsAdvance = 16;

// The kRoleScalars block of a chain in the epsilon launch (kNT threads): the
// lambda decision of a chain without the F model (nothing reads lambda between
// here and the next UpdateP); then wave 0 runs the alpha update (it reads only
// Q, whose update closed before UpdateZ, and alpha), waves 1-3 sum thirds of
// the lnP(D) tile partials; then wave 0 adds lambda and alpha (both final for
// the sweep) and the lnP accumulators into the posterior sums (output.c:50-101,
// UpdateSums :654-707).  F and its sum follow in the tail.
__device__ void scalars_block(const ChainDesc &C, const ModelArgs &m, uint32_t rep, int flags)
{
    __shared__ double s_like[3];
    const int wid = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    if (C.lam_mode == kLamP) {   // lambda, from the log sums of the new P
        lambda_final_block(C, m, rep, C.lam_part, C.l_blocks);
    }
    if (wid == 0 && (flags & kEpsAlpha) && m.inferalpha) {
        alpha_wave(C, m, rep);
    }
    if (wid >= 1 && (flags & kEpsLike)) {
        const double v = like_share(C, wid - 1, 3);
        if (lane == 0) {
            s_like[wid - 1] = v;
        }
    }
    __syncthreads();
    if (!(flags & kEpsCollect) || wid != 0) {
        return;
    }
    if (lane < m.K) {
        g(C.lambdaSum)[lane] += g(C.lambda)[lane];
        g(C.alphaSum)[lane] += g(C.alpha)[lane];
    }
    if ((flags & kEpsLike) && lane == 0) {
        const double like = (s_like[0] + s_like[1]) + s_like[2];
        SGPU_G double *like3 = g(C.like3);
        like3[0] = like;
        like3[1] += like;
        like3[2] += like * like;
    }
}

// ---------------------------------------------------------------------------
// Epsilon (NonIndependenceUpdateEpsilon, UpdateEpsilon.c:71-145).  One thread
// per (locus, pop) term; a block holds whole loci (K * (256 / K) threads) so
// the K terms of a locus meet in LDS and are summed there in the reference's
// order.  Fused: SumEpsilon += eps (output.c:684-687) with kEpsCollect, and
// with kEpsFst the allele terms of the Fst log-ratio over the updated eps
// (FlikeFreqs(new) - FlikeFreqs(old), UpdateFst.c:22-51): thread (l, k) adds
// pop k's terms at locus l to its proposal group (group k, or group 0 with
// ONEFST), taking the old-F lgamma values and the log P of the two alleles
// the epsilon move touched from the move itself (the same arguments, hence the
// same values), and the block writes one partial per group.  The proposals are
// drawn here as in the Fst decision (same stream, same values).
// Blocks of other roles ride in the same launch: kRoleFold blocks fold this
// sweep's ancestry-count partials for the next sweep's Gibbs update, and one
// kRoleScalars block per chain runs alpha, the lnP(D) sum and the scalar sums.
// ---------------------------------------------------------------------------
// Pairwise trees over a block's loci (entries j * K + k for j < lpb), in place:
// a[k] and b[k] end as the sums over j.
__device__ __forceinline__ void loci_tree(double *a, double *b, int j, int K, int lpb)
{
    for (int st = 1; st < lpb; st *= 2) {
        if (j < lpb && j % (2 * st) == 0 && j + st < lpb) {
            a[threadIdx.x] += a[threadIdx.x + st * K];
            b[threadIdx.x] += b[threadIdx.x + st * K];
        }
        __syncthreads();
    }
}

// The log-eps sums of the epsilon pass's blocks (F-model lambda) for the eps
// the chain holds now (after initialisation or a host write of eps): the same
// loci per block, terms and tree as the epsilon pass, so the same values.
__global__ __launch_bounds__(kNT) void k_eps_logsum(const ChainDesc *__restrict__ cds,
                                                     const BlockRef *__restrict__ map)
{
    SGPU_BLOCK_PROLOGUE();
    __shared__ double s_a[kNT], s_b[kNT];
    const int K = m.K;
    const int lpb = kNT / K;
    const int j = threadIdx.x / K, k = threadIdx.x % K;
    const int l = bx * lpb + j;
    const bool live = j < lpb && l < m.L;
    const int na = live ? g(C.nall)[l] : 0;
    double lg = 0.0;
    if (live && k == 0 && na > 1) {
        const SGPU_G double *e = g(C.eps) + (size_t)l * m.A;
        for (int aa = 0; aa < na; aa++) {
            lg += log_d(e[aa]);
        }
    }
    s_a[threadIdx.x] = 0.0;
    s_b[threadIdx.x] = lg;
    __syncthreads();
    loci_tree(s_a, s_b, j, K, lpb);
    if (threadIdx.x == 0) {
        g(C.lam_epart)[bx] = s_b[0];
    }
}

__device__ void eps_launch_block(const ChainDesc *__restrict__ cds, const BlockRef *__restrict__ map,
                                 uint32_t rep, int flags)
{
    SGPU_BLOCK_PROLOGUE();
    if (br_.role == kRoleFold) {
        fold_rows_block(C, m, bx, (int)br_.by);
        return;
    }
    if (br_.role == kRoleScalars) {
        scalars_block(C, m, rep, flags);
        return;
    }
    const SGPU_G int32_t *nall = g(C.nall);
    const SGPU_G double *P = g(C.P);
    const SGPU_G double *F = g(C.F);
    const SGPU_G double *lambda = g(C.lambda);
    SGPU_G double *eps = g(C.eps);
    SGPU_G double *epsSum = g(C.epsSum);
    __shared__ double s_t[kNT];         // each (locus, pop)'s part of the move's log-ratio
    __shared__ double s_f[kNT];         // Fst terms of each (locus, pop)
    __shared__ double s_g[kNT];         // log-eps sums of the loci (F-model lambda)
    __shared__ double s_newf[kMaxK];    // this sweep's Fst proposals
    const int K = m.K;
    const int lpb = kNT / K;                       // loci per block
    const int j = threadIdx.x / K, k = threadIdx.x % K;
    const int l = bx * lpb + j;
    const bool live = j < lpb && l < m.L;
    const int na_l = nall[live ? l : 0];   // loaded with the pairs below, no branch
    const bool fst = (flags & kEpsFst) != 0;   // block-uniform
    SGPU_G double *e = eps + (size_t)(live ? l : 0) * m.A;
    const SGPU_G double *pk = P + ((size_t)(live ? l : 0) * K + (live ? k : 0)) * m.A;
    // two-allele data: both alleles' eps and frequencies are loaded with the
    // locus count, before the draw picks which is a1 (16-byte aligned pairs)
    const bool two = m.A == 2;   // uniform
    double2 ev{0.0, 0.0}, pv{0.0, 0.0};
    if (two) {
        ev = *(const SGPU_G double2 *)e;
        pv = *(const SGPU_G double2 *)pk;
    }
    const double Fk = F[live ? k : 0];
    // the locus's move uniforms do not depend on the loads: formed while they
    // are in flight (only a locus with two or more alleles uses them)
    const u32x4 w = philox4x32_10(rep, (uint32_t)l, 0u, PH_EPS << 24, m.seed, m.chain);
    const int na = live ? na_l : 0;
    int a1 = 0, a2 = 0;
    double d = 0.0, e1 = 0.0, e2 = 0.0;
    bool inrange = false;
    if (fst && (int)threadIdx.x < (m.onefst ? 1 : K)) {
        s_newf[threadIdx.x] = fst_proposal(m, F[threadIdx.x], threadIdx.x, rep, nullptr);
    }
    if (na > 1) {
        a1 = (int)(uz(w.x) * (double)na);
        a2 = (int)(uz(w.y) * (double)(na - 1));
        if (a2 >= a1) {
            a2 += 1;
        }
        d = 0.0 + uz(w.z) * (m.invsqrtn - 0.0);
        e1 = two ? (a1 ? ev.y : ev.x) : e[a1];
        e2 = two ? (a2 ? ev.y : ev.x) : e[a2];
        inrange = (e1 + d < 1.0) && (e2 - d > 0.0);
    }
    const double frac = (1.0 - Fk) / Fk;
    double lg0 = 0.0, lg1 = 0.0, lg2 = 0.0, lg3 = 0.0, lp1 = 0.0, lp2 = 0.0, tk = 0.0;
    if (inrange) {
        lg0 = lgamma_lanczos(frac * e1);
        lg1 = lgamma_lanczos(frac * e2);
        lg2 = lgamma_lanczos(frac * (e1 + d));
        lg3 = lgamma_lanczos(frac * (e2 - d));
        lp1 = log_d(two ? (a1 ? pv.y : pv.x) : pk[a1]);
        lp2 = log_d(two ? (a2 ? pv.y : pv.x) : pk[a2]);
        // pop k's six terms in the reference's order within a pop
        tk = ((((lg0 + lg1) - lg2) - lg3) + frac * d * lp1) - frac * d * lp2;
    }
    s_t[threadIdx.x] = tk;
    __syncthreads();
    // every thread of the locus adds the K pop parts in pop order and decides:
    // the same sum and the same uniform, so the same decision in each
    bool acc = false;
    if (inrange) {
        const int t0 = (int)threadIdx.x - k;
        double sum = 0.0;
        for (int kk = 0; kk < K; kk++) {
            sum += s_t[t0 + kk];
        }
        const double lam = lambda[0];
        if (lam != 1.0) {
            // log(pow(x, lam - 1)) of UpdateEpsilon.c:117-118 as (lam - 1) log x
            // (equal but for rounding, and no pow)
            sum += ((double)lam - 1.0) * log_d((e1 + d) * (e2 - d) / (e1) / (e2));
        }
        acc = uo(w.w) < exp(sum);
    }
    if (live && k == 0) {
        if (acc) {
            e[a1] = e1 + d;
            e[a2] = e2 - d;
        }
        if (flags & kEpsCollect) {
            for (int aa = 0; aa < na; aa++) {
                const double ev = (acc && aa == a1) ? e1 + d : ((acc && aa == a2) ? e2 - d : e[aa]);
                epsSum[(size_t)l * m.A + aa] += ev;
            }
        }
    }
    // F-model lambda: the log-eps sum of the block's loci as this pass leaves
    // eps, for the next sweep's lambda decision
    const bool lamsum = C.lam_mode == kLamEps;   // block-uniform
    if (!fst && !lamsum) {
        return;
    }
    double lg = 0.0;
    if (lamsum && live && k == 0 && na > 1) {
        for (int aa = 0; aa < na; aa++) {
            lg += log_d(e[aa]);
        }
    }
    double v = 0.0;
    const int grp = m.onefst ? 0 : k;
    const double newf = fst ? s_newf[grp] : 0.0;
    if (live && na > 0 && newf > 0.0 && newf < 1.0) {
        const double oldf = F[grp];
        const double nfr = (1.0 - newf) / newf;
        const double ofr = (1.0 - oldf) / oldf;
        const bool reuse = inrange && ofr == frac;   // the move's lgamma arguments are ours
        const SGPU_G double *pl = P + ((size_t)l * K + k) * m.A;
        for (int aa = 0; aa < na; aa++) {
            double ep, lp, lo;
            if (reuse && aa == a1) {
                ep = acc ? e1 + d : e1;
                lp = lp1;
                lo = acc ? lg2 : lg0;
            } else if (reuse && aa == a2) {
                ep = acc ? e2 - d : e2;
                lp = lp2;
                lo = acc ? lg3 : lg1;
            } else {
                ep = (acc && aa == a1) ? e1 + d : ((acc && aa == a2) ? e2 - d : e[aa]);
                lp = log_d(pl[aa]);
                lo = lgamma_lanczos(ofr * ep);
            }
            v += (nfr * ep * lp - lgamma_lanczos(nfr * ep)) - (ofr * ep * lp - lo);
        }
    }
    // per group, a pairwise tree over the block's loci
    s_f[threadIdx.x] = v;
    s_g[threadIdx.x] = lg;
    __syncthreads();
    loci_tree(s_f, s_g, j, K, lpb);
    if (lamsum && threadIdx.x == 0) {
        g(C.lam_epart)[bx] = s_g[0];
    }
    if (!fst) {
        return;
    }
    SGPU_G double *part = g(C.fst_part);
    if (m.onefst) {
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int kk = 0; kk < K; kk++) {
                t += s_f[kk];
            }
            part[bx] = t;
        }
    } else if ((int)threadIdx.x < K) {
        part[(size_t)bx * K + threadIdx.x] = s_f[threadIdx.x];
    }
}

// Closing a sweep (kEpsAdvance, when no Fst decision follows): the block that
// finishes last (counted in *done, which it resets) advances the device sweep
// index; every block has read the index by then.
__device__ __forceinline__ void advance_when_last(uint32_t *repd, uint32_t *done, uint32_t rep)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t prev =
            __hip_atomic_fetch_add(done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (prev + 1u == gridDim.x) {
            __hip_atomic_store(done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(repd, rep + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

__global__ __launch_bounds__(kNT) void k_update_eps(const ChainDesc *__restrict__ cds,
                                                     const BlockRef *__restrict__ map,
                                                     uint32_t *__restrict__ repd,
                                                     uint32_t *__restrict__ done, int flags)
{
    const uint32_t rep = *repd;   // sweep index, read on the device (graph replays)
    eps_launch_block(cds, map, rep, flags);
    if (flags & kEpsAdvance) {
        advance_when_last(repd, done, rep);
    }
}

// ---------------------------------------------------------------------------
// Fst (UpdateFst.c:142-204).  The per-locus allele terms of
// FlikeFreqs(new) - FlikeFreqs(old) (UpdateFst.c:22-51) come in block partials,
// per proposal group: from the epsilon pass of the sweep (above), or, called
// alone, from k_fst_partial (grid: locus blocks x proposal groups, proposal
// drawn once per block).  The decision (one lane per group) adds the prior
// ratio (FPriorDiff, :10-18) and the per-locus lgamma(frac) terms.  ONEFST:
// one proposal, likelihood summed over all pops.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kNT) void k_fst_partial(const ChainDesc *__restrict__ cds,
                                                      const BlockRef *__restrict__ map,
                                                      const uint32_t *__restrict__ repp)
{
    SGPU_BLOCK_PROLOGUE();
    const uint32_t rep = *repp;   // sweep index, read on the device (graph replays)
    const SGPU_G int32_t *nall = g(C.nall);
    const SGPU_G double *eps = g(C.eps);
    const SGPU_G double *P = g(C.P);
    const SGPU_G double *F = g(C.F);
    SGPU_G double *part = g(C.fst_part);
    __shared__ double s_red[kNT / 64];
    __shared__ double s_f[2];
    const int numpops1 = m.onefst ? 1 : m.K;
    const int p1 = (int)br_.by;   // the proposal group of this block
    if (threadIdx.x == 0) {
        s_f[0] = F[p1];
        s_f[1] = fst_proposal(m, s_f[0], p1, rep, nullptr);
    }
    __syncthreads();
    const double oldf = s_f[0], newf = s_f[1];
    const int l = bx * kNT + threadIdx.x;
    const int na = (l < m.L) ? nall[l] : 0;
    double v = 0.0;
    if (newf > 0.0 && newf < 1.0 && na > 0) {
        const double nfr = (1.0 - newf) / newf;
        const double ofr = (1.0 - oldf) / oldf;
        const int p2end = m.onefst ? m.K : p1 + 1;
        const SGPU_G double *el = eps + (size_t)l * m.A;
        for (int p2 = p1; p2 < p2end; p2++) {
            const SGPU_G double *pl = P + ((size_t)l * m.K + p2) * m.A;
            // the next allele's two values are loaded one iteration ahead
            double ep_n = el[0], p_n = pl[0];
            for (int aa = 0; aa < na; aa++) {
                const double ep = ep_n, pv = p_n;
                if (aa + 1 < na) {
                    ep_n = el[aa + 1];
                    p_n = pl[aa + 1];
                }
                const double lp = log_d(pv);
                v += (nfr * ep * lp - lgamma_lanczos(nfr * ep)) -
                     (ofr * ep * lp - lgamma_lanczos(ofr * ep));
            }
        }
    }
    const double t = block_sum<kNT>(v, s_red);
    if (threadIdx.x == 0) {
        part[(size_t)bx * numpops1 + p1] = t;
    }
}

// The decision of proposal group p1 given its summed partials (lane p1 of a
// wave); with collect, FstSum += F for the group's pops (UpdateSums,
// output.c:654-707), from the values this lane itself settles.
__device__ __forceinline__ void fst_decide(const ChainDesc &C, const ModelArgs &m, uint32_t rep,
                                           int p1, double stat, bool collect)
{
    SGPU_G double *F = g(C.F);
    const double mu = m.fpriormean, sd = m.fpriorsd;
    uint32_t uw;
    const double oldf = F[p1];
    const double newf = fst_proposal(m, oldf, p1, rep, &uw);
    bool acc = false;
    if (newf > 0.0 && newf < 1.0) {
        double lpd = ((mu * mu / (sd * sd) - 1) * log_d(newf / oldf) + (oldf - newf) * mu / (sd * sd));
        const double nfr = (1.0 - newf) / newf;
        const double ofr = (1.0 - oldf) / oldf;
        const int ngroup = m.onefst ? m.K : 1;
        lpd += (double)ngroup * (double)C.nloci_nonzero * (lgamma_lanczos(nfr) - lgamma_lanczos(ofr));
        lpd += stat;
        acc = (lpd >= 0.0) || (uo(uw) < exp(lpd));
    }
    const int p2end = m.onefst ? m.K : p1 + 1;
    for (int p2 = p1; p2 < p2end; p2++) {
        const double f = acc ? newf : F[p2];
        if (acc) {
            F[p2] = newf;
        }
        if (collect) {
            g(C.FstSum)[p2] += f;
        }
    }
}

__device__ void fst_final_wave(const ChainDesc &C, const ModelArgs &m, uint32_t rep, int nparts)
{
    const int numpops1 = m.onefst ? 1 : m.K;
    // the proposal groups are independent (each reads and writes its own F), so
    // lane p1 decides group p1 with the group's sum of block partials
    const int p1 = (int)(threadIdx.x & 63);
    double sums[kMaxK + 1];
    wave_col_sums(g(C.fst_part), nparts, numpops1, sums);
    double stat = 0.0;
#pragma unroll
    for (int q = 0; q < kMaxK; q++) {
        stat = p1 == q ? sums[q] : stat;
    }
    if (p1 < numpops1) {
        fst_decide(C, m, rep, p1, stat, false);
    }
}

__global__ __launch_bounds__(kNT) void k_fst_finalize(const ChainDesc *__restrict__ cds,
                                                       const BlockRef *__restrict__ map,
                                                       const uint32_t *__restrict__ repp)
{
    SGPU_BLOCK_PROLOGUE();
    if (threadIdx.x < 64) {
        fst_final_wave(C, m, *repp, C.l_blocks);
    }
}

// ---------------------------------------------------------------------------
// Data-collection pass called alone (output.c:50-101, UpdateSums :654-707):
// scalar sums and lnP(D) from the UpdateZ partials (one block per chain).
// ---------------------------------------------------------------------------
// lnP(D) of the last UpdateZ is the sum of its tile partials, split in two
// halves summed by two waves (lane 0 of each), the halves then added in order.
__device__ __forceinline__ double like_half(const ChainDesc &C, int half)
{
    const int n = C.n_it * C.n_lt, h = (n + 1) / 2;
    return half ? wave_strided_sum(g(C.likepart) + h, n - h, 1, 0)
                : wave_strided_sum(g(C.likepart), h, 1, 0);
}

// The scalar posterior sums and the lnP accumulators, by the lanes of one
// wave: lane k adds pop k's lambda, alpha and F (independent loads, no chain
// of dependent read-modify-writes), lane 0 the lnP accumulators.
__device__ __forceinline__ void collect_scalars(const ChainDesc &C, const ModelArgs &m,
                                                bool with_like, double like)
{
    const int k = (int)(threadIdx.x & 63);
    if (k < m.K) {
        g(C.lambdaSum)[k] += g(C.lambda)[k];
        g(C.alphaSum)[k] += g(C.alpha)[k];
        if (m.freqscorr) {
            g(C.FstSum)[k] += g(C.F)[k];
        }
    }
    if (with_like && k == 0) {
        SGPU_G double *like3 = g(C.like3);
        like3[0] = like;
        like3[1] += like;
        like3[2] += like * like;
    }
}

__global__ __launch_bounds__(128) void k_collect(const ChainDesc *__restrict__ cds,
                                                 const BlockRef *__restrict__ map, int with_like)
{
    SGPU_BLOCK_PROLOGUE();
    __shared__ double s_half[2];
    const int wid = (int)(threadIdx.x >> 6);
    if (with_like) {
        const double v = like_half(C, wid);
        if ((threadIdx.x & 63) == 0) {
            s_half[wid] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        collect_scalars(C, m, with_like != 0, with_like ? s_half[0] + s_half[1] : 0.0);
    }
}

// The pass that closes a sweep, one block of kTailNT threads per chain with
// the F model: the Fst decision from the epsilon pass's partials (all 16 waves
// sum them) and, when collecting, FstSum.  With kTailAdvance the block that
// finishes last (counted in *done, which it resets) advances the device sweep
// index: every block of the sweep has used the index by then, so a graph
// replay may end here.
constexpr int kTailNT = 1024;
constexpr int kTailFst = 1, kTailCollect = 2, kTailAdvance = 4;

__global__ __launch_bounds__(kTailNT) void k_sweep_tail(const ChainDesc *__restrict__ cds,
                                                         const BlockRef *__restrict__ map,
                                                         uint32_t *__restrict__ repd,
                                                         uint32_t *__restrict__ done, int flags)
{
    SGPU_BLOCK_PROLOGUE();
    const uint32_t rep = *repd;
    __shared__ double s_part[kTailNT / 64];
    __shared__ double s_stat[kMaxK + 1];
    if ((flags & kTailFst) && m.freqscorr) {
        const int numpops1 = m.onefst ? 1 : m.K;
        block_col_sums<kTailNT>(g(C.fst_part), C.e_blocks, numpops1, s_part, s_stat);
        if ((int)threadIdx.x < numpops1) {
            fst_decide(C, m, rep, (int)threadIdx.x, s_stat[threadIdx.x], (flags & kTailCollect) != 0);
        }
    }
    if (flags & kTailAdvance) {
        advance_when_last(repd, done, rep);
    }
}

// Sum the per-tile count partials into full arrays (for state reads).
__global__ void k_reduce_counts(const uint16_t *__restrict__ part, int nparts, size_t n,
                                int32_t *__restrict__ out)
{
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) {
        return;
    }
    int32_t s = 0;
    for (int t = 0; t < nparts; t++) {
        s += part[(size_t)t * n + j];
    }
    out[j] = s;
}

// Allele counts per locus for InitFreqPriors (Init.c:197-227), one wavefront
// per locus (four per block), lanes striding over the word rows; eps = lambda
// without the F model.  The counts are exact integers: per-lane registers for
// 2-bit words (popcounts of the code bits), LDS bins for byte words.
__global__ __launch_bounds__(kNT) void k_init_eps(const ChainDesc *__restrict__ cds,
                                                   const BlockRef *__restrict__ map,
                                                   const double *__restrict__ lambda0s)
{
    SGPU_BLOCK_PROLOGUE();
    const SGPU_G uint32_t *geno = g(C.geno);
    const SGPU_G int32_t *nall = g(C.nall);
    SGPU_G double *eps = g(C.eps);
    const double lambda0 = lambda0s[br_.chain];
    __shared__ uint32_t s_bins[kNT / 64][kMaxAlleles];
    const int wid = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int l = bx * (kNT / 64) + wid;
    const bool live = l < m.L;   // wave-uniform
    const int na = live ? nall[l] : 0;
    uint32_t *bins = s_bins[wid];
    for (int aa = lane; aa < m.A; aa += 64) {
        bins[aa] = 0u;
    }
    __syncthreads();
    double total = 0.0, c0 = 0.0, c1 = 0.0;
    if (live && m.freqscorr) {
        if (C.fmt == 1) {   // words [G][Lpad]: 16 individuals x 2 bits, code 3 = missing
            const int G = (m.N + 15) / 16;
            uint32_t nvalid = 0, ncode = 0;   // valid individuals, sum of their codes
            for (int gr = lane; gr < G; gr += 64) {
                const uint32_t word = geno[(size_t)gr * C.Lpad + l];
                const int nreal = min(16, m.N - gr * 16);
                const uint32_t livebits =
                    nreal >= 16 ? 0x55555555u : (((1u << (2 * nreal)) - 1u) & 0x55555555u);
                const uint32_t lo = word & 0x55555555u, hi = (word >> 1) & 0x55555555u;
                const uint32_t valid = ~(lo & hi) & livebits;
                nvalid += (uint32_t)__popc(valid);
                ncode += (uint32_t)__popc(lo & valid) + 2u * (uint32_t)__popc(hi & valid);
            }
            // wave_sum leaves the sum in lane 0: broadcast it to the lanes that write
            const double v = __shfl(wave_sum((double)nvalid), 0, 64);
            const double cs = __shfl(wave_sum((double)ncode), 0, 64);
            total = 2.0 * v;
            c0 = 2.0 * v - cs;   // code c holds 2 - c copies of allele 0 and c of allele 1
            c1 = cs;
        } else {            // words [L][Gpad]: 4 copies of individuals 2h, 2h+1
            const int G = (m.N + 1) / 2;
            uint32_t n = 0;
            for (int h = lane; h < G; h += 64) {
                const uint32_t word = geno[(size_t)l * C.Gpad + h];
                for (int c = 0; c < 4 && h * 2 + (c >> 1) < m.N; c++) {
                    const int x = (int)((word >> (8 * c)) & 0xffu);
                    if (x != 0xff) {
                        n++;
                        if (x < na) {
                            atomicAdd(&bins[x], 1u);
                        }
                    }
                }
            }
            total = __shfl(wave_sum((double)n), 0, 64);
        }
    }
    __syncthreads();
    if (!live) {
        return;
    }
    SGPU_G double *e = eps + (size_t)l * m.A;
    for (int aa = lane; aa < m.A; aa += 64) {
        double v = 0.0;
        if (aa < na) {
            if (!m.freqscorr) {
                v = lambda0;
            } else {
                const double cnt = C.fmt == 1 ? (aa == 0 ? c0 : c1) : (double)bins[aa];
                v = (lambda0 + cnt) / ((double)na * lambda0 + total);
            }
        }
        e[aa] = v;
    }
}

}  // namespace sgpu
