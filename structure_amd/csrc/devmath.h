// Device-side random streams and special functions for the admixture sweep.
//
// Random stream contract (identical on every kernel): key = (seed, chain),
// counter = (c0, c1, c2, (phase << 24) | attempt).  Phases and index
// assignment are listed in sgpu_internal.h.  Philox4x32-10 follows Salmon et
// al., SC'11; it replaces the reference's MWC64X device streams
// (Kernels/randGen.cl, Kernels/mwc64x) and its host rand()/randomArr
// (ran.c:128-161, randGen.c:23-36).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef SGPU_PHILOX_MULHI
#define SGPU_PHILOX_MULHI 0
#endif

namespace sgpu {

struct u32x4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                               uint32_t c3, uint32_t k0, uint32_t k1)
{
#if SGPU_PHILOX_MULHI
    // separate low and high 32-bit products (v_mul_lo_u32 / v_mul_hi_u32)
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint32_t lo0 = 0xD2511F53u * c0;
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c1 = lo1;
        c3 = lo0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
#else
    // 32x32->64 products written as one 64-bit multiply each (v_mad_u64_u32);
    // the three-input XORs as one v_bitop3_b32 each (truth table 0x96)
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
#endif
    return u32x4{c0, c1, c2, c3};
}

// The same Philox4x32-10 block when c0, c2, c3 and the key are wave-uniform and
// only c1 varies across lanes (UpdateZ: sweep, individual pair and phase per
// wave, locus per lane).  In the first rounds some products and XOR operands are
// uniform; grouping the uniform operands lets them combine on the scalar unit,
// where a three-input XOR with two scalar operands would first copy one into a
// vector register.  The words are those of philox4x32_10 bit for bit.
__device__ __forceinline__ u32x4 philox4x32_10_lane1(uint32_t c0, uint32_t c1, uint32_t c2,
                                                     uint32_t c3, uint32_t k0, uint32_t k1)
{
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    // round 1: both products uniform; only c1 is per lane
    uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
    uint32_t a0 = c1 ^ ((uint32_t)(p1 >> 32) ^ k0);            // lane
    uint32_t a2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;             // uniform
    uint32_t a1 = (uint32_t)p1, a3 = (uint32_t)p0;            // uniform
    k0 += W0;
    k1 += W1;
    // round 2: p0 per lane, p1 uniform
    p0 = (uint64_t)M0 * a0;
    p1 = (uint64_t)M1 * a2;
    uint32_t b0 = (uint32_t)(p1 >> 32) ^ a1 ^ k0;             // uniform
    uint32_t b2 = (uint32_t)(p0 >> 32) ^ (a3 ^ k1);           // lane
    uint32_t b1 = (uint32_t)p1, b3 = (uint32_t)p0;            // uniform, lane
    k0 += W0;
    k1 += W1;
    // round 3: p0 uniform, p1 per lane
    p0 = (uint64_t)M0 * b0;
    p1 = (uint64_t)M1 * b2;
    uint32_t d0 = (uint32_t)(p1 >> 32) ^ (b1 ^ k0);           // lane
    uint32_t d2 = b3 ^ ((uint32_t)(p0 >> 32) ^ k1);           // lane
    uint32_t d1 = (uint32_t)p1, d3 = (uint32_t)p0;            // lane, uniform
    k0 += W0;
    k1 += W1;
    // round 4: both products per lane, c3 still uniform
    p0 = (uint64_t)M0 * d0;
    p1 = (uint64_t)M1 * d2;
    c0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), d1, k0, 0x96);
    c2 = (uint32_t)(p0 >> 32) ^ (d3 ^ k1);
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    k0 += W0;
    k1 += W1;
    // rounds 5..10: everything per lane
#pragma unroll
    for (int r = 4; r < 10; r++) {
        const uint64_t q0 = (uint64_t)M0 * c0;
        const uint64_t q1 = (uint64_t)M1 * c2;
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(q1 >> 32), c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(q0 >> 32), c3, k1, 0x96);
        c1 = (uint32_t)q1;
        c3 = (uint32_t)q0;
        c0 = n0;
        c2 = n2;
        k0 += W0;
        k1 += W1;
    }
    return u32x4{c0, c1, c2, c3};
}

// The same Philox4x32-10 block when only c2 varies across lanes (byte-format
// UpdateZ: sweep, locus and phase per wave, individual pair per lane); as
// philox4x32_10_lane1, the uniform operands of the first rounds are grouped.
__device__ __forceinline__ u32x4 philox4x32_10_lane2(uint32_t c0, uint32_t c1, uint32_t c2,
                                                     uint32_t c3, uint32_t k0, uint32_t k1)
{
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    // round 1: p0 uniform, p1 per lane
    uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
    uint32_t a0 = (uint32_t)(p1 >> 32) ^ (c1 ^ k0);           // lane
    uint32_t a2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;             // uniform
    uint32_t a1 = (uint32_t)p1, a3 = (uint32_t)p0;            // lane, uniform
    k0 += W0;
    k1 += W1;
    // round 2: p0 per lane, p1 uniform
    p0 = (uint64_t)M0 * a0;
    p1 = (uint64_t)M1 * a2;
    uint32_t b0 = a1 ^ ((uint32_t)(p1 >> 32) ^ k0);           // lane
    uint32_t b2 = (uint32_t)(p0 >> 32) ^ (a3 ^ k1);           // lane
    uint32_t b1 = (uint32_t)p1, b3 = (uint32_t)p0;            // uniform, lane
    k0 += W0;
    k1 += W1;
    // round 3: both products per lane, c1 uniform
    p0 = (uint64_t)M0 * b0;
    p1 = (uint64_t)M1 * b2;
    c0 = (uint32_t)(p1 >> 32) ^ (b1 ^ k0);
    c2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), b3, k1, 0x96);
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    k0 += W0;
    k1 += W1;
    // rounds 4..10: everything per lane
#pragma unroll
    for (int r = 3; r < 10; r++) {
        const uint64_t q0 = (uint64_t)M0 * c0;
        const uint64_t q1 = (uint64_t)M1 * c2;
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(q1 >> 32), c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(q0 >> 32), c3, k1, 0x96);
        c1 = (uint32_t)q1;
        c3 = (uint32_t)q0;
        c0 = n0;
        c2 = n2;
        k0 += W0;
        k1 += W1;
    }
    return u32x4{c0, c1, c2, c3};
}

__device__ __forceinline__ uint32_t word_of(const u32x4 &w, int j)
{
    return j == 0 ? w.x : (j == 1 ? w.y : (j == 2 ? w.z : w.w));
}

// uniform in [0,1), exact: categorical draws, RandomInteger, RandomReal (ran.c:128-147)
__device__ __forceinline__ double uz(uint32_t w) { return (double)w * 0x1p-32; }
// uniform in (0,1), exact: rnd() semantics (ran.c:152-161)
__device__ __forceinline__ double uo(uint32_t w) { return ((double)w + 0.5) * 0x1p-32; }

// ln x for the sampler and the Metropolis log-ratios.  x = m 2^e with m in
// [sqrt(1/2), sqrt(2)), f = m - 1 (exact) and s = f / (2 + f), so that
// ln(1 + f) = 2 atanh(s) = f - s (f - R(s^2)), R(z) = sum_{j>=1} 2 z^j / (2j + 1)
// (the Taylor series of 2 atanh(s) / s - 2).  |s| <= 0.1716, so ten terms
// leave a truncation error below 2^-58 of the result.  ln 2 is split into a
// 32-bit head (e * head exact for every exponent) and a tail.  The reciprocal
// of 2 + f in [1.29, 2.42] is the hardware estimate refined by two Newton
// steps.  Error about 1 ulp; 0 -> -inf, +inf -> +inf, negative or NaN -> NaN.
// About 35 instructions instead of the library log's double-double
// evaluation (over 100).  log_pos is the same evaluation for finite x > 0
// (subnormals included) without the special cases.
__device__ __forceinline__ double log_pos(double x)
{
    double m = __builtin_amdgcn_frexp_mant(x);
    int e = __builtin_amdgcn_frexp_exp(x);
    if (m < 0.7071067811865476) {
        m = m + m;
        e -= 1;
    }
    const double f = m - 1.0;
    const double den = 2.0 + f;
    double y = __builtin_amdgcn_rcp(den);
    y = fma(y, fma(-den, y, 1.0), y);
    y = fma(y, fma(-den, y, 1.0), y);
    const double s = f * y;
    const double z = s * s;
    double R = 0.09523809523809523;   // 2/21
    R = fma(R, z, 0.10526315789473684);
    R = fma(R, z, 0.11764705882352941);
    R = fma(R, z, 0.13333333333333333);
    R = fma(R, z, 0.15384615384615385);
    R = fma(R, z, 0.18181818181818182);
    R = fma(R, z, 0.2222222222222222);
    R = fma(R, z, 0.2857142857142857);
    R = fma(R, z, 0.4);
    R = fma(R, z, 0.6666666666666666);   // 2/3
    R = R * z;
    const double lm = fma(-s, f - R, f);
    const double de = (double)e;
    return fma(de, 0.6931471803691238, fma(de, 1.9082149292705877e-10, lm));
}

__device__ __forceinline__ double log_d(double x)
{
    const double r = log_pos(x);
    if (x > 0.0 && x < __builtin_inf()) {
        return r;
    }
    return x == 0.0 ? -__builtin_inf() : (x == __builtin_inf() ? x : __builtin_nan(""));
}

// sin a and cos a for 0 <= a <= 8 (the Box-Muller angle).  a = q pi/2 + r
// with q = rint(2a/pi) and r from a two-part pi/2 (Cody-Waite; the head
// product is exact under the fused multiply-add, so r keeps its relative
// accuracy next to the zeros); cos r and sin r by their Taylor series through
// r^16 and r^17 (|r| <= pi/4: truncation below 2^-60), quadrant by q mod 4:
// cos a = (c, -s, -c, s)[q mod 4], sin a = (s, c, -s, -c)[q mod 4].
// Error about 1 ulp.
__device__ __forceinline__ void sincos_d(double a, double *sn, double *cs)
{
    const double q = rint(a * 0.6366197723675814);
    double r = fma(-q, 1.5707963267948966, a);
    r = fma(-q, 6.123233995736766e-17, r);
    const int n = (int)q & 3;
    const double z = r * r;
    double c = 4.779477332387385e-14;   // 1/16!
    c = fma(c, z, -1.1470745597729725e-11);
    c = fma(c, z, 2.08767569878681e-09);
    c = fma(c, z, -2.755731922398589e-07);
    c = fma(c, z, 2.48015873015873e-05);
    c = fma(c, z, -0.001388888888888889);
    c = fma(c, z, 0.041666666666666664);
    c = fma(c, z, -0.5);
    c = fma(c, z, 1.0);
    double t = 2.8114572543455206e-15;   // 1/17!
    t = fma(t, z, -7.647163731819816e-13);
    t = fma(t, z, 1.6059043836821613e-10);
    t = fma(t, z, -2.505210838544172e-08);
    t = fma(t, z, 2.7557319223985893e-06);
    t = fma(t, z, -0.0001984126984126984);
    t = fma(t, z, 0.008333333333333333);
    t = fma(t, z, -0.16666666666666666);
    t = fma(t * z, r, r);
    const double cv = (n & 1) ? t : c;
    const double sv = (n & 1) ? c : t;
    *cs = (n == 1 || n == 2) ? -cv : cv;
    *sn = (n >= 2) ? -sv : sv;
}

__device__ __forceinline__ double cos_d(double a)
{
    double sn, cs;
    sincos_d(a, &sn, &cs);
    return cs;
}

__device__ __forceinline__ double box_muller(uint32_t w0, uint32_t w1)
{
    double r = sqrt(-2.0 * log_pos(uo(w0)));   // uo() lies in (0, 1)
    return r * cos_d(6.283185307179586 * uo(w1));
}

__device__ __forceinline__ double normal_draw(uint32_t seed, uint32_t chain, uint32_t c0,
                                              uint32_t c1, uint32_t c2, uint32_t phase)
{
    u32x4 w = philox4x32_10(c0, c1, c2, phase << 24, seed, chain);
    return box_muller(w.x, w.y);
}

constexpr uint32_t kGammaMaxBlocks = 128;        // 256 attempts
constexpr uint32_t kGammaBoostBlock = 0xFFFFFFu;   // attempt field of the shape < 1 boost

// One Marsaglia-Tsang attempt with normal x and acceptance uniform u.
__device__ __forceinline__ bool mt_attempt(double x, double u, double d, double c, double *g)
{
    double v = 1.0 + c * x;
    if (v <= 0.0) {
        return false;
    }
    v = v * v * v;
    const double x2 = x * x;
    // u in (0, 1) and v > 0 here: the logs need no special cases
    if (u < 1.0 - 0.0331 * (x2 * x2) || log_pos(u) < 0.5 * x2 + d * (1.0 - v + log_pos(v))) {
        *g = d * v;
        return true;
    }
    return false;
}

// The two attempts of Philox block 0 (see gamma_draw): the Box-Muller pair and
// its two acceptance uniforms.  They do not depend on the shape, so a caller
// can form them while the shape's inputs are still being loaded.
struct GammaFirst {
    double x0, x1, u0, u1;
};

__device__ __forceinline__ GammaFirst gamma_first(uint32_t seed, uint32_t chain, uint32_t c0,
                                                  uint32_t c1, uint32_t c2, uint32_t phase)
{
    const u32x4 w = philox4x32_10(c0, c1, c2, phase << 24, seed, chain);
    const double rad = sqrt(-2.0 * log_pos(uo(w.x)));
    double sn, cs;
    sincos_d(6.283185307179586 * uo(w.y), &sn, &cs);
    return {rad * cs, rad * sn, uo(w.z), uo(w.w)};
}

// Gamma(a, 1), Marsaglia & Tsang (ACM TOMS 26(3), 2000); a < 1 through
// Gamma(a+1) * U^(1/a).  Philox block j (attempt field j) serves two
// attempts: words 0,1 give a Box-Muller pair, the cos normal tried with the
// uniform of word 2, then the sin normal with the uniform of word 3; the boost
// uniform for a < 1 is word 0 of block kGammaBoostBlock.  FP64 throughout:
// with shapes in the thousands (allele / ancestry counts) the acceptance test
// d*(1 - v + ln v) cancels catastrophically in FP32.  Block 0's attempts come
// in f (gamma_first of the same counters), the later blocks are drawn here.
__device__ __forceinline__ double gamma_draw_from(double a, const GammaFirst &f, uint32_t seed,
                                                  uint32_t chain, uint32_t c0, uint32_t c1,
                                                  uint32_t c2, uint32_t phase)
{
    if (!(a > 0.0)) {
        return 0.0;
    }
    const double aa = a < 1.0 ? a + 1.0 : a;
    const double d = aa - 1.0 / 3.0;
    const double c = 1.0 / sqrt(9.0 * d);
    double g = 0.0;
    bool done = mt_attempt(f.x0, f.u0, d, c, &g) || mt_attempt(f.x1, f.u1, d, c, &g);
    for (uint32_t j = 1; j < kGammaMaxBlocks && !done; j++) {
        const u32x4 w = philox4x32_10(c0, c1, c2, (phase << 24) | j, seed, chain);
        const double rad = sqrt(-2.0 * log_pos(uo(w.x)));
        double sn, cs;
        sincos_d(6.283185307179586 * uo(w.y), &sn, &cs);
        done = mt_attempt(rad * cs, uo(w.z), d, c, &g) || mt_attempt(rad * sn, uo(w.w), d, c, &g);
    }
    if (!done) {
        return d;   // 256 rejections: unreachable in practice (acceptance >= 0.95)
    }
    if (a < 1.0) {
        const u32x4 b = philox4x32_10(c0, c1, c2, (phase << 24) | kGammaBoostBlock, seed, chain);
        g = g * exp(log_pos(uo(b.x)) / a);
    }
    return g;
}

__device__ __forceinline__ double gamma_draw(double a, uint32_t seed, uint32_t chain,
                                             uint32_t c0, uint32_t c1, uint32_t c2,
                                             uint32_t phase)
{
    if (!(a > 0.0)) {
        return 0.0;
    }
    return gamma_draw_from(a, gamma_first(seed, chain, c0, c1, c2, phase), seed, chain, c0, c1, c2,
                           phase);
}

// Term-by-term Lanczos sum in the host routine's order (mymath.c:382-429):
// c_8 / (z + 7) first, then down to c_1 / z.  The loop stays rolled, one
// quotient at a time: the branch is all but never taken, and unrolled its eight
// independent divisions would set the register budget of every kernel that
// calls lgamma_lanczos.
__device__ __forceinline__ double lgamma_lanczos_terms(double z)
{
    double s = 0.0, den = z + 7.0;   // denominators by repeated subtraction, as the host
#pragma unroll 1
    for (int j = 7; j >= 0; j--) {
        const double c = j == 7 ? 1.659470187408462e-7
                         : j == 6 ? 9.934937113930748e-6
                         : j == 5 ? -0.1385710331296526
                         : j == 4 ? 12.50734324009056
                         : j == 3 ? -176.6150291498386
                         : j == 2 ? 771.3234287757674
                         : j == 1 ? -1259.139216722289
                                  : 676.5203681218835;
        s += c / den;
        den -= 1.0;
    }
    s += 0.9999999999995183;
    return log_pos(s) + 0.9189385332046727 - (z + 6.5) + (z - 0.5) * log_pos(z + 6.5);
}

// ln Gamma(z), z > 0: Lanczos g = 7, n = 9 with the published coefficients,
// the approximation the reference uses on the host (mymath.c:382-429).  The
// eight partial fractions c_j / (z + j - 1) are accumulated as one fraction
// n / d (n <- n e + c d with one fused multiply-add, d <- d e), so the sum
// costs one quotient instead of eight: the hardware reciprocal refined by two
// Newton steps and one residual correction of n y.  The value agrees with the
// term-by-term sum to a few ulps.  Beyond z = 1e30 (d would overflow) the
// term-by-term sum is used.
__device__ __forceinline__ double lgamma_lanczos(double z)
{
    if (z > 1e30) {
        return lgamma_lanczos_terms(z);
    }
    const double c[8] = {676.5203681218835, -1259.139216722289, 771.3234287757674,
                         -176.6150291498386, 12.50734324009056, -0.1385710331296526,
                         9.934937113930748e-6, 1.659470187408462e-7};
    double n = 0.0, d = 1.0, e = z + 7.0;
#pragma unroll
    for (int j = 7; j >= 0; j--) {   // z+7 down to z, as the host routine walks them
        n = fma(n, e, c[j] * d);
        d = d * e;
        e -= 1.0;
    }
    double y = __builtin_amdgcn_rcp(d);
    y = fma(y, fma(-d, y, 1.0), y);
    y = fma(y, fma(-d, y, 1.0), y);
    double q = n * y;
    q = fma(fma(-d, q, n), y, q);
    const double s = q + 0.9999999999995183;
    return log_pos(s) + 0.9189385332046727 - (z + 6.5) + (z - 0.5) * log_pos(z + 6.5);
}

// Wave-level (64 lanes) and block-level FP64 sums with a fixed reduction tree.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_down(v, off, 64);
    }
    return v;
}

// Block sum over nw waves (blockDim.x = 64 nw): the same tree as block_sum<64 nw>.
__device__ __forceinline__ double block_sum_waves(double v, double *lds /* nw */, int nw)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        lds[wid] = v;
    }
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < nw; i++) {
            t += lds[i];
        }
    }
    __syncthreads();
    return t;   // valid in thread 0
}

template <int NT>
__device__ __forceinline__ double block_sum(double v, double *lds /* NT/64 */)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        lds[wid] = v;
    }
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < NT / 64; i++) {
            t += lds[i];
        }
    }
    __syncthreads();
    return t;   // valid in thread 0
}

}  // namespace sgpu
