/*
 * structure_align.h -- label alignment of replicate runs and chain-level
 * summaries (no counterpart in the reference program, which runs one chain;
 * users align replicates with CLUMPP afterwards).
 */
#ifndef STRUCTURE_ALIGN_H
#define STRUCTURE_ALIGN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Minimum-cost assignment (Hungarian method) of the n x n row-major cost
 * matrix: row_to_col[i] is the column assigned to row i.  n <= 64. */
int shost_assign(const double *cost, int n, int32_t *row_to_col);

/* Permutation aligning q [N][K] to qref [N][K]: cluster a of qref matches
 * cluster perm[a] of q, minimising sum_i,a (qref[i][a] - q[i][perm[a]])^2. */
int shost_align_q(const double *qref, const double *q, int32_t N, int32_t K, int32_t *perm);

/* out[r][a][:] = x[r][perm[a]][:] for x laid out [rows][K][inner] (Q: inner
 * 1; P: rows = loci, inner = alleles). */
int shost_permute_clusters(const double *x, int32_t rows, int32_t K, int32_t inner,
                           const int32_t *perm, double *out);

/* R replicates of one K: qbars [R][N][K] posterior-mean Q, lnp [R] estimated
 * ln P(D).  Aligns every replicate to replicate 0 and returns the mean and
 * standard deviation of the aligned Q, the mean ln P(D) and its Monte Carlo
 * standard error (sd / sqrt(R)); perms [R][K] (may be NULL) receives the
 * permutations. */
int shost_replicate_summary(const double *qbars, const double *lnp, int32_t R, int32_t N,
                            int32_t K, double *qmean, double *qsd, double *lnp_mean,
                            double *lnp_se, int32_t *perms);

#ifdef __cplusplus
}
#endif
#endif
