// Fused UpdateZ kernels: categorical draw of every allele copy's source
// population (UpdateZ.c:13-56, PickAnOptionDiscrete mymath.c:105-129), the
// per-locus allele-by-population histogram (GetNumFromPops, UpdateP.c:60-99),
// the per-individual ancestry counts (GetNumLociPop, UpdateQ.c:17-35) and,
// optionally, the lnP(D) terms of CalcLike (CalcLike.c:75-133).  Z itself is
// never written to HBM (a debug copy can be requested).
//
// Mapping: one locus per lane, 256 loci per block (4 wavefronts); a block
// walks the individuals of its individual tile.  Each lane keeps its locus's
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

// Per-individual count accumulation: mask of lanes whose copy went to pop k.
__device__ __forceinline__ void add_ind_counts(uint32_t *s_cnt, int slot, int K, int lane,
                                               const int z0, const int z1, const bool v0,
                                               const bool v1)
{
    for (int k = 0; k < K; k++) {
        const unsigned long long m0 = __ballot(v0 && z0 == k);
        const unsigned long long m1 = __ballot(v1 && z1 == k);
        const uint32_t n = (uint32_t)(__popcll(m0) + __popcll(m1));
        if (lane == 0 && n) {
            atomicAdd(&s_cnt[slot * K + k], n);
        }
    }
}

// Running product of clamped terms with exponent bookkeeping (CalcLikeInd's
// running product, CalcLike.c:113-126, without the periodic logs).
struct LikeAcc {
    double prod = 1.0;
    int expo = 0;
    __device__ __forceinline__ void mul(double t) { prod *= (t > kSqrtUnder) ? t : kSqrtUnder; }
    __device__ __forceinline__ void renorm()
    {
        int e;
        prod = frexp(prod, &e);
        expo += e;
    }
    __device__ __forceinline__ double log_value() const { return log(prod) + (double)expo * kLn2; }
};

// ---------------------------------------------------------------------------
// Biallelic 2-bit path.  K is a template parameter (registers, unrolled dots).
// ---------------------------------------------------------------------------
template <int K, bool INIT, bool LIKE>
__global__ __launch_bounds__(kZLoci) void k_update_z_snp2(ZArgs a)
{
    extern __shared__ uint32_t s_cnt[];   // [tile individuals][K]
    const int lane = threadIdx.x & 63;
    const int l = blockIdx.x * kZLoci + threadIdx.x;   // < Lpad by construction
    const bool lok = l < a.L;
    const int row0 = blockIdx.y * a.rows_per_tile;
    const int row1 = min(row0 + a.rows_per_tile, a.G);
    const int ibase = row0 * 16;
    const int ntile = min((row1 - row0) * 16, a.N - ibase);

    for (int t = threadIdx.x; t < ntile * K; t += kZLoci) {
        s_cnt[t] = 0;
    }
    __syncthreads();

    double P0[K], P1[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        P0[k] = 0.0;
        P1[k] = 0.0;
        if (!INIT && lok) {
            P0[k] = a.P[((size_t)l * K + k) * a.A + 0];
            P1[k] = a.P[((size_t)l * K + k) * a.A + 1];
        }
    }
    int cnt0[K], cnt1[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        cnt0[k] = 0;
        cnt1[k] = 0;
    }
    LikeAcc acc;

    for (int g = row0; g < row1; g++) {
        const uint32_t word = a.geno[(size_t)g * a.Lpad + l];
        for (int jp = 0; jp < 8; jp++) {
            const int i0 = g * 16 + 2 * jp;
            if (i0 >= a.N) {
                break;
            }
            const u32x4 w = philox4x32_10(a.rep, (uint32_t)l, (uint32_t)(i0 >> 1),
                                          (INIT ? PH_INITZ : PH_Z) << 24, a.seed, a.chain);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int i = i0 + h;
                if (i >= a.N) {
                    break;
                }
                const uint32_t code = (word >> (2 * (2 * jp + h))) & 3u;
                const bool valid = code != 3u;
                const int x0 = code == 2u;
                const int x1 = code >= 1u;
                const uint32_t wa = h ? w.z : w.x;
                const uint32_t wb = h ? w.w : w.y;
                int z0 = 0, z1 = 0;
                if (INIT) {
                    z0 = (int)(((uint64_t)wa * (uint32_t)K) >> 32);
                    z1 = (int)(((uint64_t)wb * (uint32_t)K) >> 32);
                } else {
                    double pre0[K], pre1[K];
                    double s0 = 0.0, s1 = 0.0;
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const double q = a.Q[(size_t)i * K + k];
                        const double c0 = q * P0[k];
                        const double c1 = q * P1[k];
                        s0 = s0 + c0;
                        s1 = s1 + c1;
                        pre0[k] = s0;
                        pre1[k] = s1;
                    }
                    const double S0 = x0 ? s1 : s0;
                    const double S1 = x1 ? s1 : s0;
                    const double r0 = uz(wa) * S0;
                    const double r1 = uz(wb) * S1;
#pragma unroll
                    for (int k = 0; k < K - 1; k++) {
                        z0 += ((x0 ? pre1[k] : pre0[k]) < r0);
                        z1 += ((x1 ? pre1[k] : pre0[k]) < r1);
                    }
                    if (LIKE && valid) {
                        acc.mul(S0);
                        acc.mul(S1);
                    }
                }
                if (valid) {
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        cnt0[k] += (z0 == k && !x0) + (z1 == k && !x1);
                        cnt1[k] += (z0 == k && x0) + (z1 == k && x1);
                    }
                }
                add_ind_counts(s_cnt, i - ibase, K, lane, z0, z1, valid, valid);
                if (a.zout && lok) {
                    a.zout[((size_t)i * 2 + 0) * a.L + l] = valid ? (int8_t)z0 : (int8_t)-9;
                    a.zout[((size_t)i * 2 + 1) * a.L + l] = valid ? (int8_t)z1 : (int8_t)-9;
                }
            }
            if (LIKE) {
                acc.renorm();
            }
        }
    }

    if (lok) {
        int32_t *dst = a.nApart + ((size_t)blockIdx.y * a.L + l) * K * a.A;
#pragma unroll
        for (int k = 0; k < K; k++) {
            dst[k * a.A + 0] = cnt0[k];
            dst[k * a.A + 1] = cnt1[k];
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < ntile * K; t += kZLoci) {
        a.nQpart[((size_t)blockIdx.x * a.N + ibase) * K + t] = (int32_t)s_cnt[t];
    }
    if (LIKE) {
        __shared__ double s_red[kZLoci / 64];
        const double v = lok ? acc.log_value() : 0.0;
        const double tot = block_sum<kZLoci>(v, s_red);
        if (threadIdx.x == 0) {
            a.likepart[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = tot;
        }
    }
}

// ---------------------------------------------------------------------------
// Byte-per-copy path (any allele count): runtime K <= kMaxK.  A word holds the
// two copies of two individuals, so one Philox block serves one word.  The
// lane owns its locus's partial histogram row and updates it in place.
// ---------------------------------------------------------------------------
template <bool INIT, bool LIKE>
__global__ __launch_bounds__(kZLoci) void k_update_z_byte(ZArgs a)
{
    extern __shared__ uint32_t s_cnt[];
    const int lane = threadIdx.x & 63;
    const int K = a.K;
    const int l = blockIdx.x * kZLoci + threadIdx.x;
    const bool lok = l < a.L;
    const int row0 = blockIdx.y * a.rows_per_tile;
    const int row1 = min(row0 + a.rows_per_tile, a.G);
    const int ibase = row0 * 2;
    const int ntile = min((row1 - row0) * 2, a.N - ibase);

    for (int t = threadIdx.x; t < ntile * K; t += kZLoci) {
        s_cnt[t] = 0;
    }
    int32_t *hist = a.nApart + ((size_t)blockIdx.y * a.L + (lok ? l : 0)) * K * a.A;
    if (lok) {
        for (int t = 0; t < K * a.A; t++) {
            hist[t] = 0;
        }
    }
    __syncthreads();
    const double *Pl = a.P + (size_t)(lok ? l : 0) * K * a.A;
    LikeAcc acc;

    for (int g = row0; g < row1; g++) {
        const uint32_t word = a.geno[(size_t)g * a.Lpad + l];
        const int i0 = g * 2;
        const u32x4 w = philox4x32_10(a.rep, (uint32_t)l, (uint32_t)(i0 >> 1),
                                      (INIT ? PH_INITZ : PH_Z) << 24, a.seed, a.chain);
        for (int h = 0; h < 2; h++) {
            const int i = i0 + h;
            if (i >= a.N) {
                break;
            }
            const int x0 = (int)((word >> (16 * h)) & 0xffu);
            const int x1 = (int)((word >> (16 * h + 8)) & 0xffu);
            const bool v0 = lok && x0 != 0xff;
            const bool v1 = lok && x1 != 0xff;
            // uniforms go to the non-missing copies in order (randGen.c:23-36)
            const uint32_t wa = h ? w.z : w.x;
            const uint32_t wb1 = v0 ? (h ? w.w : w.y) : wa;
            int z0 = 0, z1 = 0;
            if (INIT) {
                z0 = (int)(((uint64_t)wa * (uint32_t)K) >> 32);
                z1 = (int)(((uint64_t)wb1 * (uint32_t)K) >> 32);
            } else {
                double q[kMaxK];
                for (int k = 0; k < K; k++) {
                    q[k] = a.Q[(size_t)i * K + k];
                }
                if (v0) {
                    double s = 0.0;
                    for (int k = 0; k < K; k++) {
                        s = s + q[k] * Pl[k * a.A + x0];
                    }
                    const double r = uz(wa) * s;
                    double run = 0.0;
                    for (int k = 0; k < K - 1; k++) {
                        run = run + q[k] * Pl[k * a.A + x0];
                        z0 += run < r;
                    }
                    if (LIKE) {
                        acc.mul(s);
                    }
                }
                if (v1) {
                    double s = 0.0;
                    for (int k = 0; k < K; k++) {
                        s = s + q[k] * Pl[k * a.A + x1];
                    }
                    const double r = uz(wb1) * s;
                    double run = 0.0;
                    for (int k = 0; k < K - 1; k++) {
                        run = run + q[k] * Pl[k * a.A + x1];
                        z1 += run < r;
                    }
                    if (LIKE) {
                        acc.mul(s);
                    }
                }
            }
            if (v0) {
                hist[z0 * a.A + x0] += 1;
            }
            if (v1) {
                hist[z1 * a.A + x1] += 1;
            }
            add_ind_counts(s_cnt, i - ibase, K, lane, z0, z1, v0, v1);
            if (a.zout && lok) {
                a.zout[((size_t)i * 2 + 0) * a.L + l] = v0 ? (int8_t)z0 : (int8_t)-9;
                a.zout[((size_t)i * 2 + 1) * a.L + l] = v1 ? (int8_t)z1 : (int8_t)-9;
            }
        }
        if (LIKE) {
            acc.renorm();
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < ntile * K; t += kZLoci) {
        a.nQpart[((size_t)blockIdx.x * a.N + ibase) * K + t] = (int32_t)s_cnt[t];
    }
    if (LIKE) {
        __shared__ double s_red[kZLoci / 64];
        const double v = lok ? acc.log_value() : 0.0;
        const double tot = block_sum<kZLoci>(v, s_red);
        if (threadIdx.x == 0) {
            a.likepart[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = tot;
        }
    }
}

}  // namespace sgpu
