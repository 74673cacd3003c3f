// Internal definitions shared by the sgpu kernels and the C-ABI layer.
#pragma once

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
constexpr int kMaxK = 16;          // largest K the kernels take
constexpr int kMaxAlleles = 255;   // byte format

struct ZArgs {
    const uint32_t *geno;      // packed words
    const double *Q;           // [N][K]
    const double *P;           // [L][K][A]
    int32_t *nApart;           // [n_it][L][K][A]
    int32_t *nQpart;           // [n_lt][N][K]
    double *likepart;          // [n_it][n_lt]
    int8_t *zout;              // [N][2][L] or null
    int N, L, K, A, Lpad, G;   // G = number of genotype word rows
    int rows_per_tile;         // word rows per individual tile
    int inds_per_row;          // 16 (SNP2) or 2 (BYTE)
    uint32_t seed, chain;
};

}  // namespace sgpu
