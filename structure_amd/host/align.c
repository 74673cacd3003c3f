/* Label alignment of replicate runs and chain-level summaries.
 *
 * Mixture labels are arbitrary: cluster 2 of one replicate may be cluster 5
 * of another.  Replicates of one K are aligned to a reference replicate by
 * the permutation that minimises the squared distance between their
 * posterior-mean Q matrices (an assignment problem, solved exactly by the
 * Hungarian method in O(K^3)); the aligned Q are then averaged, as CLUMPP
 * does for STRUCTURE output.  Not part of the reference program. */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sgpu.h"
#include "../../include/structure_align.h"

/* Minimum-cost assignment of rows to columns of the n x n cost matrix c
 * (Kuhn-Munkres with potentials, shortest augmenting paths).  row_to_col[i]
 * receives the column of row i. */
int shost_assign(const double *c, int n, int32_t *row_to_col)
{
    if (n < 1 || n > 64) {
        return SGPU_ERR_ARG;
    }
    /* 1-based arrays with a virtual column 0 */
    double u[65], v[65], minv[65];
    int p[65], way[65], used[65];
    for (int j = 0; j <= n; j++) {
        u[j] = v[j] = 0.0;
        p[j] = 0;
        way[j] = 0;
    }
    for (int i = 1; i <= n; i++) {
        p[0] = i;
        int j0 = 0;
        for (int j = 0; j <= n; j++) {
            minv[j] = DBL_MAX;
            used[j] = 0;
        }
        do {
            used[j0] = 1;
            const int i0 = p[j0];
            double delta = DBL_MAX;
            int j1 = 0;
            for (int j = 1; j <= n; j++) {
                if (used[j]) {
                    continue;
                }
                const double cur = c[(size_t)(i0 - 1) * n + (j - 1)] - u[i0] - v[j];
                if (cur < minv[j]) {
                    minv[j] = cur;
                    way[j] = j0;
                }
                if (minv[j] < delta) {
                    delta = minv[j];
                    j1 = j;
                }
            }
            for (int j = 0; j <= n; j++) {
                if (used[j]) {
                    u[p[j]] += delta;
                    v[j] -= delta;
                } else {
                    minv[j] -= delta;
                }
            }
            j0 = j1;
        } while (p[j0] != 0);
        do {   /* augment along the path */
            const int j1 = way[j0];
            p[j0] = p[j1];
            j0 = j1;
        } while (j0);
    }
    for (int j = 1; j <= n; j++) {
        row_to_col[p[j] - 1] = j - 1;
    }
    return SGPU_OK;
}

int shost_align_q(const double *qref, const double *q, int32_t N, int32_t K, int32_t *perm)
{
    if (!qref || !q || !perm || N < 1 || K < 1 || K > 64) {
        return SGPU_ERR_ARG;
    }
    /* cost[a][b] = sum_i (qref[i][a] - q[i][b])^2 */
    double *cost = (double *)calloc((size_t)K * K, sizeof(double));
    if (!cost) {
        return SGPU_ERR_NOMEM;
    }
    for (int i = 0; i < N; i++) {
        for (int a = 0; a < K; a++) {
            for (int b = 0; b < K; b++) {
                const double d = qref[(size_t)i * K + a] - q[(size_t)i * K + b];
                cost[(size_t)a * K + b] += d * d;
            }
        }
    }
    const int rc = shost_assign(cost, K, perm);
    free(cost);
    return rc;
}

int shost_permute_clusters(const double *x, int32_t rows, int32_t K, int32_t inner,
                           const int32_t *perm, double *out)
{
    if (!x || !out || !perm || rows < 0 || K < 1 || inner < 1) {
        return SGPU_ERR_ARG;
    }
    for (int r = 0; r < rows; r++) {
        for (int a = 0; a < K; a++) {
            memcpy(out + ((size_t)r * K + a) * inner, x + ((size_t)r * K + perm[a]) * inner,
                   sizeof(double) * (size_t)inner);
        }
    }
    return SGPU_OK;
}

int shost_replicate_summary(const double *qbars, const double *lnp, int32_t R, int32_t N,
                            int32_t K, double *qmean, double *qsd, double *lnp_mean,
                            double *lnp_se, int32_t *perms)
{
    if (!qbars || !lnp || R < 1 || N < 1 || K < 1 || K > 64 || !qmean || !lnp_mean ||
        !lnp_se) {
        return SGPU_ERR_ARG;
    }
    const size_t nk = (size_t)N * K;
    double *aligned = (double *)malloc(sizeof(double) * nk);
    double *sq = (double *)calloc(nk, sizeof(double));
    int32_t perm[64];
    if (!aligned || !sq) {
        free(aligned);
        free(sq);
        return SGPU_ERR_NOMEM;
    }
    memset(qmean, 0, sizeof(double) * nk);
    int rc = SGPU_OK;
    for (int r = 0; r < R && rc == SGPU_OK; r++) {
        const double *q = qbars + (size_t)r * nk;
        if (r == 0) {
            for (int k = 0; k < K; k++) {
                perm[k] = k;
            }
        } else if ((rc = shost_align_q(qbars, q, N, K, perm))) {
            break;
        }
        if (perms) {
            memcpy(perms + (size_t)r * K, perm, sizeof(int32_t) * K);
        }
        shost_permute_clusters(q, N, K, 1, perm, aligned);
        for (size_t j = 0; j < nk; j++) {
            qmean[j] += aligned[j];
            sq[j] += aligned[j] * aligned[j];
        }
    }
    if (rc == SGPU_OK) {
        for (size_t j = 0; j < nk; j++) {
            qmean[j] /= R;
            if (qsd) {
                const double var = R > 1 ? (sq[j] - R * qmean[j] * qmean[j]) / (R - 1) : 0.0;
                qsd[j] = var > 0.0 ? sqrt(var) : 0.0;
            }
        }
        double s = 0.0, s2 = 0.0;
        for (int r = 0; r < R; r++) {
            s += lnp[r];
            s2 += lnp[r] * lnp[r];
        }
        *lnp_mean = s / R;
        const double var = R > 1 ? (s2 - R * (*lnp_mean) * (*lnp_mean)) / (R - 1) : 0.0;
        *lnp_se = var > 0.0 ? sqrt(var / R) : 0.0;
    }
    free(aligned);
    free(sq);
    return rc;
}
