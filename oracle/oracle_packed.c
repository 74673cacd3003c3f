/*
 * CPU oracle, large-data helpers -- TEST INFRASTRUCTURE ONLY.
 *
 * The same functions as oracle.c restated over the 2-bit genotype words
 * (16 individuals x 2 bits per uint32, words [ceil(N/16)][Lpad], code 0 =
 * (0,0), 1 = (0,1), 2 = (1,1), 3 = missing) and spread over host threads,
 * for checks at data sizes where the [N][2][L] int8 layout of oracle.c would
 * not fit in memory (N = 20000, L = 500000: 20 GB as int8, 2.5 GB packed).
 *
 * orc_calc_like_snp2 is CalcLike (CalcLike.c:387-426) -> CalcLikeInd
 * (:75-133) with the individuals split over threads.  Each individual's
 * running product visits (line, locus) in the reference's order and the
 * per-individual totals are added in individual order at the end, so the
 * result equals orc_calc_like on the unpacked genotypes bit for bit.
 *
 * Only tests/ and bench.py's measurement legs use this file, as the checker.
 */
#define _POSIX_C_SOURCE 200809L
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define ORC_SQRTUNDER_P 1e-50   /* sqrt(UNDERFLO), CalcLike.c:99 */

static int orc_threads(int want)
{
    if (want > 0) {
        return want;
    }
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    if (n < 1) {
        n = 1;
    }
    return n > 64 ? 64 : (int)n;
}

/* Allele of copy c (0 or 1) of a 2-bit code, -1 when missing. */
static inline int orc_snp2_allele(uint32_t code, int c)
{
    if (code == 3u) {
        return -1;
    }
    return c == 0 ? (code == 2u) : (code != 0u);
}

typedef struct {
    const orc_model *m;
    const uint32_t *words;
    int Lpad;
    const double *Q, *P;
    double *ind;        /* [N] per-individual log likelihood */
    int row0, row1;     /* word rows of this thread */
} like_job;

static void *like_worker(void *arg)
{
    like_job *j = (like_job *)arg;
    const int N = j->m->N, L = j->m->L, K = j->m->K, A = j->m->A;
    for (int g = j->row0; g < j->row1; g++) {
        const uint32_t *row = j->words + (size_t)g * j->Lpad;
        const int nreal = N - 16 * g < 16 ? N - 16 * g : 16;
        double run[16], tot[16];
        for (int t = 0; t < 16; t++) {
            run[t] = 1.0;
            tot[t] = 0.0;
        }
        /* CalcLikeInd: line-major (all loci of line 0, then line 1) */
        for (int c = 0; c < 2; c++) {
            for (int l = 0; l < L; l++) {
                const uint32_t w = row[l];
                const double *Pl = j->P + (size_t)l * K * A;
                for (int t = 0; t < nreal; t++) {
                    const int x = orc_snp2_allele((w >> (2 * t)) & 3u, c);
                    if (x < 0) {
                        continue;
                    }
                    const double *q = j->Q + (size_t)(16 * g + t) * K;
                    double term = 0.0;
                    for (int k = 0; k < K; k++) {
                        term += q[k] * Pl[k * A + x];
                    }
                    run[t] *= (term > ORC_SQRTUNDER_P) ? term : ORC_SQRTUNDER_P;
                    if (run[t] < ORC_SQRTUNDER_P) {
                        tot[t] += log(run[t]);
                        run[t] = 1.0;
                    }
                }
            }
        }
        for (int t = 0; t < nreal; t++) {
            j->ind[16 * g + t] = tot[t] + log(run[t]);
        }
    }
    return NULL;
}

double orc_calc_like_snp2(const orc_model *m, const uint32_t *words, int Lpad, const double *Q,
                          const double *P, int nthreads)
{
    const int N = m->N;
    const int G = (N + 15) / 16;
    double *ind = (double *)calloc((size_t)N, sizeof(double));
    int nt = orc_threads(nthreads);
    if (nt > G) {
        nt = G;
    }
    pthread_t *th = (pthread_t *)calloc((size_t)nt, sizeof(pthread_t));
    like_job *jobs = (like_job *)calloc((size_t)nt, sizeof(like_job));
    for (int t = 0; t < nt; t++) {
        jobs[t] = (like_job){m, words, Lpad, Q, P, ind, (int)((long long)G * t / nt),
                             (int)((long long)G * (t + 1) / nt)};
        pthread_create(&th[t], NULL, like_worker, &jobs[t]);
    }
    for (int t = 0; t < nt; t++) {
        pthread_join(th[t], NULL);
    }
    double loglike = 0.0;
    for (int i = 0; i < N; i++) {   /* CalcLike's sum, individual order */
        loglike += ind[i];
    }
    free(ind);
    free(th);
    free(jobs);
    return loglike;
}

typedef struct {
    int N, L, Lpad;
    const uint32_t *words;
    int64_t *loc_valid, *loc_a1, *ind_valid;
    int l0, l1;         /* loci of this thread */
    int row0, row1;     /* word rows of this thread */
} count_job;

static void *count_worker(void *arg)
{
    count_job *j = (count_job *)arg;
    const int G = (j->N + 15) / 16;
    for (int l = j->l0; l < j->l1; l++) {
        int64_t v = 0, a1 = 0;
        for (int g = 0; g < G; g++) {
            const uint32_t w = j->words[(size_t)g * j->Lpad + l];
            const int nreal = j->N - 16 * g < 16 ? j->N - 16 * g : 16;
            for (int t = 0; t < nreal; t++) {
                const uint32_t code = (w >> (2 * t)) & 3u;
                if (code != 3u) {
                    v += 2;
                    a1 += code;   /* code c carries c copies of allele 1 */
                }
            }
        }
        j->loc_valid[l] = v;
        j->loc_a1[l] = a1;
    }
    for (int g = j->row0; g < j->row1; g++) {
        const int nreal = j->N - 16 * g < 16 ? j->N - 16 * g : 16;
        int64_t v[16] = {0};
        for (int l = 0; l < j->L; l++) {
            const uint32_t w = j->words[(size_t)g * j->Lpad + l];
            for (int t = 0; t < nreal; t++) {
                v[t] += ((w >> (2 * t)) & 3u) != 3u ? 2 : 0;
            }
        }
        for (int t = 0; t < nreal; t++) {
            j->ind_valid[16 * g + t] = v[t];
        }
    }
    return NULL;
}

void orc_snp2_counts(int N, int L, int Lpad, const uint32_t *words, int64_t *loc_valid,
                     int64_t *loc_a1, int64_t *ind_valid, int nthreads)
{
    const int G = (N + 15) / 16;
    int nt = orc_threads(nthreads);
    pthread_t *th = (pthread_t *)calloc((size_t)nt, sizeof(pthread_t));
    count_job *jobs = (count_job *)calloc((size_t)nt, sizeof(count_job));
    for (int t = 0; t < nt; t++) {
        jobs[t] = (count_job){N, L, Lpad, words, loc_valid, loc_a1, ind_valid,
                              (int)((long long)L * t / nt), (int)((long long)L * (t + 1) / nt),
                              (int)((long long)G * t / nt), (int)((long long)G * (t + 1) / nt)};
        pthread_create(&th[t], NULL, count_worker, &jobs[t]);
    }
    for (int t = 0; t < nt; t++) {
        pthread_join(th[t], NULL);
    }
    free(th);
    free(jobs);
}
