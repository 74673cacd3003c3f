// Internal definitions shared by the sgpu kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgpu {

// Random stream phases: counter c3 = (phase << 24) | attempt.
//   PH_Z      c0=rep c1=locus c2=ind>>1   word 2*(ind&1)+j, j-th non-missing copy
//   PH_INITZ  c0=0   c1=locus c2=ind>>1   same words; z = floor(u*K)
//   PH_P      c0=rep c1=locus c2=(pop<<16)|allele   Marsaglia-Tsang attempts
//   PH_Q      c0=rep c1=ind   c2=pop                Marsaglia-Tsang attempts
//   PH_QMETRO c0=rep c1=ind   c2=pop (proposal) | 0xffffffff (acceptance, word 0)
//   PH_ALPHA / PH_LAMBDA / PH_FST  c0=rep c1=pop c2=0: words 0,1 normal, 2 uniform
//   PH_EPS    c0=rep c1=locus c2=0: word 0 allele 1, 1 allele 2, 2 step, 3 uniform
enum : uint32_t {
    PH_Z = 1, PH_INITZ = 2, PH_P = 3, PH_Q = 4, PH_QMETRO = 5,
    PH_ALPHA = 6, PH_LAMBDA = 7, PH_EPS = 8, PH_FST = 9
};

constexpr double kUnderflo = 1e-100;     // structure.h:6
constexpr double kSqrtUnder = 1e-50;     // CalcLike.c:99
constexpr double kLambdaMax = 10.0;      // structure.h:17
constexpr double kLambdaPropSd = 0.3;    // structure.h:18
constexpr double kLn2 = 0.6931471805599453;

constexpr int kZLoci = 256;        // loci per UpdateZ block (one locus per lane)
constexpr int kMaxK = 20;          // largest K the kernels take (MAXPOPS; UpdateZ.c:41-46)
constexpr int kMaxAlleles = 255;   // byte format
// Most UpdateZ locus tiles whose ancestry partials the Gibbs update sums itself.
constexpr int kQDirectRows = 32;

struct ModelArgs {
    int N, L, K, A;
    int freqscorr, onefst, popalphas, unifprioralpha, popspecificlambda;
    int inferalpha, inferlambda, pad0;
    int ancestdist, numboxes;      // ANCESTDIST histogram of Q (collected sweeps)
    int pfonly, pad1;              // PFROMPOPFLAGONLY: allele counts from POPFLAG=1 only
    double alphapropsd, alphamax, alphapriora, alphapriorb, fpriormean, fpriorsd;
    double invsqrtn;   // N^-1/2 from the host's pow, the epsilon step bound (UpdateEpsilon.c:80)
    uint32_t seed, chain;
};

// One chain of a batch as the kernels see it: its model, its tilings and the
// device addresses of its state.  A context keeps an array of these in device
// memory; chains may differ in N, L, K, A, data set and flags.
struct ChainDesc {
    ModelArgs m;
    int fmt;                       // 1 = 2-bit SNP words, 2 = byte per copy
    int Lpad, G;                   // padded loci per word row (2-bit), word rows
    int Gpad;                      // byte format: padded word rows, the stride between loci
    int n_lt, n_it, rows_per_tile; // UpdateZ tiles: loci tiles, individual tiles, rows per tile
    int q_direct;                  // Gibbs Q sums the n_lt ancestry partials itself (no fold)
    int zt_loci, z_chunk;          // loci per UpdateZ block; byte path: loci staged in LDS at once
    int q_blocks, l_blocks;        // cdiv(N, 256), cdiv(L, 256)
    int e_blocks;                  // epsilon-pass blocks: cdiv(L, 256 / K)
    int n_ranges, loci_per_range;  // Metropolis locus ranges
    int nhist, nloci_nonzero;      // allele-count histogram size, loci with alleles
    int p_blocks;                  // UpdateP blocks of this chain
    int lam_mode;                  // kLamNone .. kLamSplit
    // UpdateP work list [p_blocks][256]: (l << 13 | k << 8 | a) for a < nall[l],
    // each (locus, pop) group whole within one block, ~0 = idle thread; null:
    // the dense (locus, pop, allele < A) grid
    const uint32_t *pslots;
    const uint32_t *geno;          // packed words: 2-bit [G][Lpad], byte [L][Gpad]
    const int32_t *nall;           // [L]
    const int32_t *ahist;          // (allele count, loci) pairs
    double *Q, *P, *eps, *F, *alpha, *lambda;
    // count partials of one UpdateZ tile each: at most 24576 copies per tile
    // (the LDS cap on a 2-bit tile's rows), so 16 bits hold them
    uint16_t *nApart;              // [n_it][L][K][A]
    uint16_t *nQpart;              // [n_lt][N][K]
    int32_t *nQfold;               // [16][N][K]
    int32_t *logq_rows;            // rows of logq_part written by the last Q update
    double *likepart;              // [n_it][n_lt]
    double *logq_part;             // [rows][K+1]
    double *delta_part;            // [n_ranges][N]
    double *qprop;                 // [N][K] Metropolis proposals of the current sweep
    // lambda log-sum partials [l_blocks][K]; Fst log-ratio partials: [l_blocks][K]
    // (Fst pass) or [e_blocks][K] (epsilon pass)
    double *lam_part, *fst_part;
    double *lam_epart;             // [e_blocks] log-eps sums of the last epsilon pass (kLamEps)
    double *QSum, *PSum, *epsSum, *FstSum, *alphaSum, *lambdaSum, *like3;
    int8_t *zout;                  // debug copy of Z [N][2][L] or null
    int32_t *adist;                // [N][K][numboxes] ANCESTDIST histogram, or null
    // PFROMPOPFLAGONLY (null when off): 2-bit words, per word row the bits 2t of
    // the individuals 16g+t whose POPFLAG is 1 ([G]); byte words, POPFLAG per
    // individual ([N])
    const uint32_t *pfmask;
    const int32_t *popflag;
};

#ifdef __HIP_DEVICE_COMPILE__
// Global-memory view of a device address read from a descriptor: lets the
// compiler emit global (not flat) loads and stores.
#define SGPU_G __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ SGPU_G T *g(T *p)
{
    return (SGPU_G T *)p;
}
#else
#define SGPU_G
template <class T>
__host__ __device__ inline T *g(T *p)
{
    return p;
}
#endif

// A block of a batched launch: its chain, its block coordinates within that
// chain's own grid, and its role where one launch runs several passes.
struct BlockRef {
    uint32_t chain, bx, by, role;
};

// Block roles of the launches that carry more than one pass.
enum : uint32_t {
    kRoleMain = 0,      // the launch's own pass
    kRoleLambda = 1,    // UpdateP launch: the lambda decision of an F-model chain
    kRoleFold = 2,      // epsilon launch: one block of the ancestry-count fold
    kRoleScalars = 3,   // epsilon launch: alpha, the lnP(D) sum and the scalar posterior sums
    kRoleLamPart = 4,   // Metropolis proposal launch: 256 loci of the lambda log sums of P
    kRoleQGibbs = 5,    // UpdateP launch of a Gibbs sweep: a block of the Gibbs Q update
    kRoleMetroProp = 6  // UpdateP launch of a Metropolis sweep: a block of the Q proposals
};

// Where a sweep runs a chain's lambda update (UpdateLambda.c:8-102): its log
// sums read eps (F model) or the new P, and its decision must precede the
// epsilon update, which reads lambda.
enum : int {
    kLamNone = 0,    // lambda not inferred
    // F model: the log-eps sums of the eps that each epsilon pass leaves, the
    // decision in the next sweep's UpdateP launch (nothing before the epsilon
    // update reads lambda or writes eps)
    kLamEps = 1,
    // no F model: log sums of the new P after UpdateP (the Metropolis proposal
    // launch, or a launch of their own in Gibbs sweeps), decision in the epsilon launch
    kLamP = 2,
    kLamSplit = 3    // F model with pop-specific lambda (log sums of the new P): own launches
};

}  // namespace sgpu
