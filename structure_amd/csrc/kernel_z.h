// Fused UpdateZ kernels: categorical draw of every allele copy's source
// population (UpdateZ.c:13-56, PickAnOptionDiscrete mymath.c:105-129), the
// per-locus allele-by-population histogram (GetNumFromPops, UpdateP.c:60-99),
// the per-individual ancestry counts (GetNumLociPop, UpdateQ.c:17-35) and,
// optionally, the lnP(D) terms of CalcLike (CalcLike.c:75-133).  Z itself is
// never written to HBM (a debug copy can be requested).
//
// Mapping of the 2-bit kernel: one locus per lane, 256 loci per block (4
// wavefronts); a block walks the individuals of its individual tile.  Each lane keeps its locus's
// P column and allele-by-population counters in registers, so the per-locus
// histogram needs no atomics; Q[i,:] is wave-uniform (scalar loads); the
// per-individual counts are wave ballots reduced in LDS.
//
// The categorical draw runs in FP64 in the reference's operation order
// (products Q*P, running sum from 0, r = u*sum, first k with r <= prefix):
// with contraction disabled the device reproduces the reference's double
// arithmetic bit for bit, so Z needs no guard band.  On gfx950 without packed
// FP32 math, FP64 add/mul issue at the FP32 rate.
#pragma once

#include "devmath.h"
#include "sgpu_internal.h"

namespace sgpu {

// Lane mask of a predicate (the condition's own mask, no VGPR round trip).
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// (bit of this lane in mask) ? v : 0, the mask staying in SGPRs.
__device__ __forceinline__ uint32_t lane_select(uint64_t mask, uint32_t v)
{
    uint32_t r;
    asm("v_cndmask_b32_e64 %0, 0, %1, %2" : "=v"(r) : "v"(v), "s"(mask));
    return r;
}

// v_writelane_b32 (the compiler's intrinsic: it moves the lane index into M0
// itself, as gfx9 encodes the instruction with a scalar data operand)
extern "C" __device__ int sgpu_writelane(int value, int lane, int old) __asm("llvm.amdgcn.writelane.i32");

// v with lane `lane` replaced by the wave-uniform `value`.
__device__ __forceinline__ uint32_t write_lane(uint32_t v, uint32_t value, int lane)
{
    return (uint32_t)sgpu_writelane((int)value, lane, (int)v);
}

// Running product of clamped terms with exponent bookkeeping (CalcLikeInd's
// running product, CalcLike.c:113-126, without the periodic logs).
// max(t, c) for the finite non-negative terms here, without the NaN
// canonicalisation the compiler adds to fmax.
// The bound c is wave-uniform (a constant): it stays in scalar registers, so the
// loop does not copy it into vector registers before each use.
__device__ __forceinline__ double max_raw(double t, double c)
{
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(t), "s"(c));
    return r;
}

struct LikeAcc {
    double prod = 1.0;
    int expo = 0;
    __device__ __forceinline__ void mul(double t) { prod *= (t > kSqrtUnder) ? t : kSqrtUnder; }
    // both copies of one individual, terms given scaled by 2^-32 (exactly: no
    // subnormals arise), a missing individual giving 0, 0 and so the factor
    // (sqrt(UNDERFLO) 2^-32)^2 without a select; log_value adds 32 to the
    // exponent per valid copy and takes that factor out once per missing individual
    __device__ __forceinline__ void mul_pair_scaled(double g0, double g1)
    {
        const double c = kSqrtUnder * 0x1p-32;
        prod *= max_raw(g0, c) * max_raw(g1, c);
    }
    __device__ __forceinline__ void renorm()
    {
        int e;
        prod = frexp(prod, &e);
        expo += e;
    }
    __device__ __forceinline__ double log_value(int extra_expo = 0, int missing = 0) const
    {
        const double v = log(prod) + (double)(expo + extra_expo) * kLn2;
        return missing ? v - (double)(2 * missing) * log(kSqrtUnder * 0x1p-32) : v;
    }
};

// ---------------------------------------------------------------------------
// Biallelic 2-bit path.  K is a template parameter (registers, unrolled dots).
//
// Per lane (locus) and individual, with the pop-k prefix sums pre_a[k] of
// allele a and r_c = u_c * S_c for copy c:  z_c > k  <=>  pre_{x_c}[k] < r_c.
// The kernel only ever forms these "z > k" predicates:
//   per-locus counters   TA[k] = #copies with z > k (low 16 bits) plus
//                        #allele-1 copies with z > k (high 16 bits), one
//                        lane-private register per k: a copy adds 1 or
//                        0x10001 by its allele, two selects and one
//                        three-input add per individual (a tile holds at most
//                        65535 copies, so the halves never carry into each other);
//   per-individual       popcount of the wave ballot of each predicate (SALU),
//                        two individuals' counts packed into one 32-bit value
//                        (16 bits each: at most 128 copies per wave, 512 per
//                        tile), written into lane jp of a collector register
//                        (v_writelane), and added into LDS once per 16 individuals.
// Counts per population are differences of consecutive "z > k" counts.
// Q[i,:] (wave uniform) is read through the constant address space, so it is
// fetched with scalar loads.
// ---------------------------------------------------------------------------
// Waves per SIMD asked of the compiler for the K-templated 2-bit UpdateZ where
// a few registers fewer than its own allocation lift the occupancy a step and
// the spills stay out of the inner loop; 1 = the compiler's choice.  Measured
// per K at N=2000, L=100000 (ms per launch, compiler's choice -> this):
// K=5 0.534 -> 0.525 (six waves), 7 0.715 -> 0.678, 10 1.027 -> 0.939, 11 1.115 ->
// 1.084, 14 1.712 -> 1.385, 15 1.910 -> 1.567, 16 2.027 -> 1.907; the same step made
// K = 4, 6, 8, 9, 12, 13, 17..20 slower (spills inside the loop).  K=5 takes a
// second step to seven waves (72 VGPRs; the 8 spilled stay in the per-row code,
// outside the pair loop): 0.5095 -> 0.5081 ms collecting, 0.474 -> 0.468 burn-in;
// K=6 likewise six waves (80 VGPRs, none spilled in the pair loop): 0.595 -> 0.593
// collecting, 0.559 -> 0.553 burn-in.  At K = 8, 12, 13 and 17 the next step spills
// inside the pair loop.
template <int K>
constexpr int z_snp2_waves()
{
    return K == 5 ? 7 : K == 6 ? 6 : K == 7 ? 5 : (K == 10 || K == 11) ? 4 : (K >= 14 && K <= 16) ? 3 : 1;
}

// Individual pairs per iteration of the 2-bit UpdateZ's pair loop: two, so that
// one pair's Philox rounds overlap the other's compares.  Measured at N=2000,
// L=100000 (ms per launch, one pair -> two): K=2 0.337 -> 0.332, 4 0.466 -> 0.460,
// 5 0.535 -> 0.529, 8 0.781 -> 0.777, 10 0.931 -> 0.926, 17 2.082 -> 2.067; K=20
// 2.537 -> 2.550.  One pair at K = 11, 15 and 16, whose tighter register budgets
// above would spill the second pair, and from K = 18 on.
template <int K>
constexpr int z_snp2_pair_unroll()
{
    return (K == 11 || K == 15 || K == 16 || K >= 18) ? 1 : 2;
}

// Burn-in sweeps (no lnP terms): each copy's r against both alleles' prefixes as
// four products instead of four 64-bit selects and two products.  The two more
// live doubles keep the register class (waves per SIMD, no new spill) at these
// K; at K = 4, 7, 10, 11 and 14-16 they would cost a wave or spill, and those K
// keep the selects.
template <int K>
constexpr bool z_snp2_burnin_products()
{
    return !(K == 4 || K == 7 || (K >= 10 && K <= 11) || (K >= 14 && K <= 16));
}

template <int K, bool INIT, bool LIKE, bool ZOUT, bool PF>
__global__ __launch_bounds__(kZLoci) __attribute__((amdgpu_waves_per_eu(z_snp2_waves<K>())))
void k_update_z_snp2(const ChainDesc *__restrict__ cds,
                                                          const BlockRef *__restrict__ map,
                                                          const uint32_t *__restrict__ repp)
{
    const BlockRef br = map[blockIdx.x];
    const ChainDesc &C = cds[br.chain];
    const uint32_t rep = *repp;   // sweep index, read on the device (graph replays)
    const int bx = (int)br.bx, by = (int)br.by;
    const int N = C.m.N, L = C.m.L, A = C.m.A, Lpad = C.Lpad, G = C.G;
    const int rows_per_tile = C.rows_per_tile, n_lt = C.n_lt;
    const uint32_t seed = C.m.seed, chain = C.m.chain;
    const SGPU_G uint32_t *geno = g(C.geno);
    // Q is not written during the launch: the constant address space lets the
    // wave-uniform Q[i,:] reads go through scalar loads
    const __attribute__((address_space(4))) double *Q =
        (const __attribute__((address_space(4))) double *)C.Q;
    const SGPU_G double *P = g(C.P);
    SGPU_G uint16_t *nApart = g(C.nApart);
    SGPU_G uint16_t *nQpart = g(C.nQpart);
    SGPU_G double *likepart = g(C.likepart);
    SGPU_G int8_t *zout = g(C.zout);
    // PFROMPOPFLAGONLY (PF): per word row, bit 2t set when individual 16g+t adds to
    // the allele counts
    const SGPU_G uint32_t *pfmask = g(C.pfmask);
    // [tile individual pair][K]: valid copies, then z > k; the pair's even individual
    // in the low 16 bits, the odd one in the high 16 bits
    extern __shared__ uint32_t s_cnt[];
    constexpr int kPairUnroll = z_snp2_pair_unroll<K>();
    const int lane = threadIdx.x & 63;
    const int l = bx * kZLoci + threadIdx.x;   // < Lpad by construction
    const bool lok = l < L;
    const int row0 = by * rows_per_tile;
    const int row1 = min(row0 + rows_per_tile, G);
    const int ibase = row0 * 16;
    const int ntile = min((row1 - row0) * 16, N - ibase);

    const int npairs = (ntile + 1) / 2;
    for (int t = threadIdx.x; t < npairs * K; t += kZLoci) {
        s_cnt[t] = 0;
    }
    __syncthreads();

    double P0[K], P1[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        P0[k] = 0.0;
        P1[k] = 0.0;
        if (!INIT && lok) {
            P0[k] = P[((size_t)l * K + k) * A + 0];
            P1[k] = P[((size_t)l * K + k) * A + 1];
        }
    }
    uint32_t TA[K];   // entries 0..K-2 used
#pragma unroll
    for (int k = 0; k < K; k++) {
        TA[k] = 0;
    }
    uint32_t Ttot = 0, A1tot = 0;   // copies (and allele-1 copies) the allele counts take
    uint32_t Tall = 0;              // every valid copy (the lnP terms)
    LikeAcc acc;

    uint32_t word = geno[(size_t)row0 * Lpad + l];
    for (int g = row0; g < row1; g++) {
        const uint32_t next = (g + 1 < row1) ? geno[(size_t)(g + 1) * Lpad + l] : 0u;
        const uint32_t pfm = PF ? pfmask[g] : 0x55555555u;   // wave-uniform
        {
            // this row's valid copies and allele-1 copies at the lane's locus, from the
            // word's bit pattern (code 3 = missing; only individuals < N count)
            const int nreal = min(16, N - g * 16);
            const uint32_t live =
                nreal >= 16 ? 0x55555555u : (((1u << (2 * nreal)) - 1u) & 0x55555555u);
            const uint32_t lo = word & 0x55555555u, hi = (word >> 1) & 0x55555555u;
            const uint32_t vall = ~(lo & hi) & live;
            const uint32_t vmask = vall & pfm;
            Tall += 2u * (uint32_t)__popc(vall);
            if (PF) {   // without the filter the allele counts take every valid copy
                Ttot += 2u * (uint32_t)__popc(vmask);
            }
            A1tot += (uint32_t)__popc(lo & ~hi & vmask) + 2u * (uint32_t)__popc(hi & ~lo & vmask);
        }
        uint32_t col[K];   // collector: lane jp holds the packed counts of pair jp of the row
#pragma unroll
        for (int k = 0; k < K; k++) {
            col[k] = 0;
        }
#pragma unroll kPairUnroll
        for (int jp = 0; jp < 8; jp++) {
            const int i0 = g * 16 + 2 * jp;
            if (i0 >= N) {
                break;
            }
            const u32x4 w = philox4x32_10_lane1(rep, (uint32_t)l, (uint32_t)(i0 >> 1),
                                                (INIT ? PH_INITZ : PH_Z) << 24, seed, chain);
            uint32_t nkp[K];   // the pair's counts, individual i0 low, i0 + 1 high
#pragma unroll
            for (int k = 0; k < K; k++) {
                nkp[k] = 0;
            }
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int i = i0 + h;
                if (i >= N) {
                    break;
                }
                const int t = 2 * jp + h;
                const uint32_t code = __builtin_amdgcn_ubfe(word, 2u * (uint32_t)t, 2u);
                const bool valid = code != 3u;
                const bool x0 = code == 2u;
                const bool x1 = code != 0u;
                const uint32_t X0 = x0 ? 0x10001u : 1u, X1 = x1 ? 0x10001u : 1u;
                // lanes whose copies go into the allele counts: all, or (PF) none when
                // the individual's POPFLAG is not 1 (its ancestry counts still take them)
                const uint64_t fm = (!PF || ((pfm >> (2 * t)) & 1u)) ? ~0ull : 0ull;
                const uint32_t wa = h ? w.z : w.x;
                const uint32_t wb = h ? w.w : w.y;
                const uint64_t vm = ballot(valid);
                uint32_t nk[K];   // this individual's counts (wave-uniform)
                nk[0] = 2u * (uint32_t)__popcll(vm);
                int z0 = 0, z1 = 0;
                if (INIT) {
                    const uint32_t u0 = (uint32_t)(((uint64_t)wa * (uint32_t)K) >> 32);
                    const uint32_t u1 = (uint32_t)(((uint64_t)wb * (uint32_t)K) >> 32);
#pragma unroll
                    for (int k = 0; k < K - 1; k++) {
                        const bool m0 = valid && u0 > (uint32_t)k;
                        const bool m1 = valid && u1 > (uint32_t)k;
                        const uint64_t b0 = ballot(m0), b1 = ballot(m1);
                        TA[k] += lane_select(b0 & fm, X0) + lane_select(b1 & fm, X1);
                        nk[k + 1] = (uint32_t)(__popcll(b0) + __popcll(b1));
                    }
                } else {
                    double pre0[K], pre1[K];
                    double s0 = 0.0, s1 = 0.0;
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const double q = Q[(size_t)i * K + k];
                        const double c0 = q * P0[k];
                        const double c1 = q * P1[k];
                        // the reference's running sum starts at 0.0; 0.0 + c == c for the
                        // non-negative products here, so the first term starts it
                        s0 = k ? s0 + c0 : c0;
                        s1 = k ? s1 + c1 : c1;
                        pre0[k] = s0;
                        pre1[k] = s1;
                    }
                    // r = (w * 2^-32) * S, formed as w * (S * 2^-32): the power-of-two
                    // scaling is exact, so both orders round the same product once.
                    // Missing copies scale S by 0 instead: r = 0 never exceeds a prefix
                    // (all >= 0), so they never count.
                    const double sc = __hiloint2double(valid ? 0x3DF00000 : 0, 0);   // 2^-32 or 0
                    const double h0 = s0 * sc, h1 = s1 * sc;
                    // r of copy c against allele a's prefixes: ra* for allele 0, rb* for 1
                    double ra0, rb0, ra1, rb1;
                    if (LIKE || !z_snp2_burnin_products<K>()) {
                        const double g0 = x0 ? h1 : h0, g1 = x1 ? h1 : h0;   // each copy's S * sc
                        if (LIKE) {
                            acc.mul_pair_scaled(g0, g1);
                        }
                        ra0 = rb0 = (double)wa * g0;
                        ra1 = rb1 = (double)wb * g1;
                    } else {
                        // in the lanes whose mask below takes a compare, its r is the
                        // same product as the selected one
                        ra0 = (double)wa * h0;
                        rb0 = (double)wa * h1;
                        ra1 = (double)wb * h0;
                        rb1 = (double)wb * h1;
                    }
                    // both allele sequences are compared and the copy's allele picks
                    // the lane mask in scalar registers
                    const uint64_t xm0 = ballot(x0), xm1 = ballot(x1);
#pragma unroll
                    for (int k = 0; k < K - 1; k++) {
                        const uint64_t a0 = ballot(pre0[k] < ra0), b0 = ballot(pre1[k] < rb0);
                        const uint64_t a1 = ballot(pre0[k] < ra1), b1 = ballot(pre1[k] < rb1);
                        const uint64_t m0 = (xm0 & b0) | (~xm0 & a0);   // z0 > k
                        const uint64_t m1 = (xm1 & b1) | (~xm1 & a1);   // z1 > k
                        TA[k] += lane_select(m0 & fm, X0) + lane_select(m1 & fm, X1);
                        nk[k + 1] = (uint32_t)(__popcll(m0) + __popcll(m1));
                        if (ZOUT) {
                            z0 += (int)((m0 >> lane) & 1u);
                            z1 += (int)((m1 >> lane) & 1u);
                        }
                    }
                }
#pragma unroll
                for (int k = 0; k < K; k++) {
                    nkp[k] |= nk[k] << (16 * h);
                }
                if (ZOUT && lok) {
                    zout[((size_t)i * 2 + 0) * L + l] = valid ? (int8_t)z0 : (int8_t)-9;
                    zout[((size_t)i * 2 + 1) * L + l] = valid ? (int8_t)z1 : (int8_t)-9;
                }
            }
            // park the pair's counts in lane jp of the collector (one VALU per value)
#pragma unroll
            for (int k = 0; k < K; k++) {
                col[k] = write_lane(col[k], nkp[k], jp);
            }
            if (LIKE) {
                acc.renorm();
            }
        }
        // add this row's per-pair counts into LDS (lanes 0..7 of each wave)
        const int p0 = (g * 16 - ibase) / 2;
        if (lane < 8 && p0 + lane < npairs) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                atomicAdd(&s_cnt[(p0 + lane) * K + k], col[k]);
            }
        }
        word = next;
    }

    if (lok) {
        SGPU_G uint16_t *dst = nApart + ((size_t)by * L + l) * K * A;
        uint32_t prevT = PF ? Ttot : Tall, prevA = A1tot;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint32_t tk = (k < K - 1) ? (TA[k] & 0xFFFFu) : 0u;
            const uint32_t ak = (k < K - 1) ? (TA[k] >> 16) : 0u;
            const uint32_t n1 = prevA - ak;
            const uint32_t nt = prevT - tk;
            dst[k * A + 0] = (uint16_t)(nt - n1);
            dst[k * A + 1] = (uint16_t)n1;
            prevT = tk;
            prevA = ak;
        }
    }
    __syncthreads();
    // convert "valid, z > 0, ..., z > K-2" into counts per population
    for (int t = threadIdx.x; t < ntile; t += kZLoci) {
        SGPU_G uint16_t *dst = nQpart + ((size_t)bx * N + ibase + t) * K;
        const uint32_t *src = s_cnt + (t >> 1) * K;
        const int sh = 16 * (t & 1);
        uint32_t prev = (src[0] >> sh) & 0xFFFFu;
        for (int k = 0; k < K; k++) {
            const uint32_t gk = (k < K - 1) ? (src[1 + k] >> sh) & 0xFFFFu : 0u;
            dst[k] = (uint16_t)(prev - gk);
            prev = gk;
        }
    }
    if (LIKE) {
        __shared__ double s_red[kZLoci / 64];
        // every valid copy's term was multiplied in scaled by 2^-32, every missing
        // individual as the clamp value squared
        const double v = lok ? acc.log_value(32 * (int)Tall, ntile - (int)(Tall / 2u)) : 0.0;
        const double tot = block_sum<kZLoci>(v, s_red);
        if (threadIdx.x == 0) {
            likepart[(size_t)by * n_lt + bx] = tot;
        }
    }
}

// ---------------------------------------------------------------------------
// Byte-per-copy path (any allele count), K templated.  One lane per genotype
// word row, i.e. per pair of individuals (2h, 2h+1), 64 to 256 rows per block;
// the block walks a tile of loci, and with the words stored locus-major a
// wavefront's rows are adjacent words of one locus.  Per locus the K x A allele frequencies are
// wave-uniform, so they are staged in LDS a chunk of loci at a time (allele
// major: a copy reads its allele's K frequencies at constant offsets), and the
// locus's allele-by-population histogram is a set of LDS bins every lane adds
// into.  Per-individual counts stay in lane registers as "z > k" counts, as in
// the 2-bit path.  One Philox block per (locus, pair): words 0,1 serve the two
// copies of individual 2h, words 2,3 those of 2h+1, the j-th non-missing copy
// taking the j-th word of its individual (randGen.c:23-36).
// ---------------------------------------------------------------------------
constexpr int kByteRows = 256;   // most word rows (individual pairs) per UpdateZ block; a batch
                                 // may use 128 or 64 (the block size)

// One allele copy: its draw, its "z > k" counts (T, V) and, when `count`, its
// bin of the locus's allele-by-population histogram (s_hl).
template <int K, bool INIT, bool LIKE, bool PF>
__device__ __forceinline__ void byte_copy(int x, bool valid, bool count, uint32_t w,
                                          const double *q, const double *s_pl, int A,
                                          uint32_t *s_hl, uint32_t *T, uint32_t &V, LikeAcc &acc,
                                          int &zout)
{
    if (!valid) {
        return;
    }
    V += 1;
    int z = 0;
    if (INIT) {
        const uint32_t u = (uint32_t)(((uint64_t)w * (uint32_t)K) >> 32);
#pragma unroll
        for (int k = 0; k < K - 1; k++) {
            const bool gt = u > (uint32_t)k;
            T[k] += gt;
            z += gt;
        }
    } else {
        // the copy's allele row of the staged frequencies: K doubles, constant offsets
        const double *px = s_pl + x * K;
        double pre[K];
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            // product and running sum in the reference's order; the sum starts at
            // 0.0, and 0.0 + c == c for the non-negative products here
            const double c = q[k] * px[k];
            s = k ? s + c : c;
            pre[k] = s;
        }
        if (LIKE) {
            acc.mul(s);
        }
        const double r = (double)w * (s * 0x1p-32);   // u * S with u = w 2^-32, exactly
#pragma unroll
        for (int k = 0; k < K - 1; k++) {
            const bool gt = pre[k] < r;   // z > k
            T[k] += gt;
            z += gt;
        }
    }
    if (!PF || count) {
        atomicAdd(&s_hl[z * A + x], 1u);
    }
    zout = z;
}

// Waves per SIMD asked for the byte-format UpdateZ (as z_snp2_waves), measured
// on microsatellites (N=2000, L=20000, 2-16 alleles; ms per launch): K=12 0.350
// -> 0.341, K=17 0.508 -> 0.452; the same step at K=5 and 8 was slower or even.
template <int K>
constexpr int z_byte_waves() { return K == 12 ? 4 : K == 17 ? 3 : 1; }

template <int K, bool INIT, bool LIKE, bool ZOUT, bool PF>
__global__ __launch_bounds__(kByteRows) __attribute__((amdgpu_waves_per_eu(z_byte_waves<K>())))
void k_update_z_byte(const ChainDesc *__restrict__ cds,
                                                             const BlockRef *__restrict__ map,
                                                             const uint32_t *__restrict__ repp)
{
    const BlockRef br = map[blockIdx.x];
    const ChainDesc &C = cds[br.chain];
    const uint32_t rep = *repp;   // sweep index, read on the device (graph replays)
    const int bx = (int)br.bx, by = (int)br.by;
    const int N = C.m.N, L = C.m.L, A = C.m.A, Gpad = C.Gpad, G = C.G;
    const int zt = C.zt_loci, chunk = C.z_chunk, n_lt = C.n_lt;
    const uint32_t seed = C.m.seed, chain = C.m.chain;
    const SGPU_G uint32_t *geno = g(C.geno);
    const SGPU_G double *Q = g(C.Q);
    const SGPU_G double *P = g(C.P);
    SGPU_G uint16_t *nApart = g(C.nApart);
    SGPU_G uint16_t *nQpart = g(C.nQpart);
    SGPU_G double *likepart = g(C.likepart);
    SGPU_G int8_t *zout = g(C.zout);
    extern __shared__ double s_dyn[];
    double *s_p = s_dyn;                                  // [chunk][A][K] frequencies
    uint32_t *s_h = (uint32_t *)(s_dyn + (size_t)chunk * K * A);   // [chunk][K][A] counts

    const int R = (int)blockDim.x;                // word rows per block (64, 128 or 256)
    const int h = by * R + threadIdx.x;           // word row of this lane
    const bool rowok = h < G;
    const int i0 = 2 * h, i1 = 2 * h + 1;
    const bool ok0 = rowok && i0 < N, ok1 = rowok && i1 < N;
    // PFROMPOPFLAGONLY (PF): only individuals with POPFLAG 1 add to the allele counts
    const SGPU_G int32_t *popflag = g(C.popflag);
    const bool cnt0 = !PF || (ok0 && popflag[i0] == 1);
    const bool cnt1 = !PF || (ok1 && popflag[i1] == 1);
    double q0[K], q1[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        q0[k] = (!INIT && ok0) ? Q[(size_t)i0 * K + k] : 0.0;
        q1[k] = (!INIT && ok1) ? Q[(size_t)i1 * K + k] : 0.0;
    }
    uint32_t T0[K], T1[K];   // entries 0..K-2: copies with z > k
#pragma unroll
    for (int k = 0; k < K; k++) {
        T0[k] = 0;
        T1[k] = 0;
    }
    uint32_t V0 = 0, V1 = 0;
    LikeAcc acc;
    const int lt0 = bx * zt, lt1 = min(lt0 + zt, L);
    const int KA = K * A;
    for (int c0 = lt0; c0 < lt1; c0 += chunk) {
        const int nc = min(chunk, lt1 - c0);
        __syncthreads();
        if (!INIT) {
            // P[l][k][a] -> s_p[l - c0][a][k], one (locus, allele) row of K per thread
            for (int r = threadIdx.x; r < nc * A; r += R) {
                const int j = r / A, a = r - j * A;
                const SGPU_G double *src = P + (size_t)(c0 + j) * KA + a;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    s_p[r * K + k] = src[k * A];
                }
            }
        }
        for (int t = threadIdx.x; t < nc * KA; t += R) {
            s_h[t] = 0u;
        }
        __syncthreads();
        // words [L][Gpad]: the lanes' pairs are adjacent words of one locus; the
        // next locus's word is loaded one iteration ahead
        uint32_t nextw = rowok ? geno[(size_t)c0 * Gpad + h] : 0xffffffffu;
        for (int j = 0; j < nc; j++) {
            const int l = c0 + j;
            const uint32_t word = nextw;
            nextw = (rowok && j + 1 < nc) ? geno[(size_t)(l + 1) * Gpad + h] : 0xffffffffu;
            if (word != 0xffffffffu) {
                const u32x4 w = philox4x32_10_lane2(rep, (uint32_t)l, (uint32_t)h,
                                              (INIT ? PH_INITZ : PH_Z) << 24, seed, chain);
                const double *s_pl = s_p + (size_t)j * KA;
                uint32_t *s_hl = s_h + (size_t)j * KA;
                const int xa = (int)(word & 0xffu), xb = (int)((word >> 8) & 0xffu);
                const int xc = (int)((word >> 16) & 0xffu), xd = (int)((word >> 24) & 0xffu);
                const bool va = ok0 && xa != 0xff, vb = ok0 && xb != 0xff;
                const bool vc = ok1 && xc != 0xff, vd = ok1 && xd != 0xff;
                int za = -9, zb = -9, zc = -9, zd = -9;
                byte_copy<K, INIT, LIKE, PF>(xa, va, cnt0, w.x, q0, s_pl, A, s_hl, T0, V0, acc, za);
                byte_copy<K, INIT, LIKE, PF>(xb, vb, cnt0, va ? w.y : w.x, q0, s_pl, A, s_hl, T0, V0,
                                         acc, zb);
                byte_copy<K, INIT, LIKE, PF>(xc, vc, cnt1, w.z, q1, s_pl, A, s_hl, T1, V1, acc, zc);
                byte_copy<K, INIT, LIKE, PF>(xd, vd, cnt1, vc ? w.w : w.z, q1, s_pl, A, s_hl, T1, V1,
                                         acc, zd);
                if (LIKE) {
                    acc.renorm();
                }
                if (ZOUT) {
                    if (ok0) {
                        zout[((size_t)i0 * 2 + 0) * L + l] = (int8_t)za;
                        zout[((size_t)i0 * 2 + 1) * L + l] = (int8_t)zb;
                    }
                    if (ok1) {
                        zout[((size_t)i1 * 2 + 0) * L + l] = (int8_t)zc;
                        zout[((size_t)i1 * 2 + 1) * L + l] = (int8_t)zd;
                    }
                }
            } else if (ZOUT && rowok) {
                if (ok0) {
                    zout[((size_t)i0 * 2 + 0) * L + l] = -9;
                    zout[((size_t)i0 * 2 + 1) * L + l] = -9;
                }
                if (ok1) {
                    zout[((size_t)i1 * 2 + 0) * L + l] = -9;
                    zout[((size_t)i1 * 2 + 1) * L + l] = -9;
                }
            }
        }
        __syncthreads();
        // this block alone covers (individual tile by, loci c0..c0+nc): plain stores
        SGPU_G uint16_t *dst = nApart + ((size_t)by * L + c0) * KA;
        for (int t = threadIdx.x; t < nc * KA; t += R) {
            dst[t] = (uint16_t)s_h[t];
        }
    }
    // counts per population from the "z > k" counts
    if (ok0) {
        SGPU_G uint16_t *d0 = nQpart + ((size_t)bx * N + i0) * K;
        uint32_t prev = V0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint32_t tk = (k < K - 1) ? T0[k] : 0u;
            d0[k] = (uint16_t)(prev - tk);
            prev = tk;
        }
    }
    if (ok1) {
        SGPU_G uint16_t *d1 = nQpart + ((size_t)bx * N + i1) * K;
        uint32_t prev = V1;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint32_t tk = (k < K - 1) ? T1[k] : 0u;
            d1[k] = (uint16_t)(prev - tk);
            prev = tk;
        }
    }
    if (LIKE) {
        __shared__ double s_red[kByteRows / 64];
        const double v = (ok0 || ok1) ? acc.log_value() : 0.0;
        const double tot = block_sum_waves(v, s_red, R >> 6);
        if (threadIdx.x == 0) {
            likepart[(size_t)by * n_lt + bx] = tot;
        }
    }
}

}  // namespace sgpu
