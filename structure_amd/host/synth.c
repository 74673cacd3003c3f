/*
 * Seeded synthetic genotype generators (include/sgpu_synth.h).
 * Random numbers: xoshiro256** (Blackman & Vigna, 2018) seeded by
 * splitmix64; gamma by Marsaglia & Tsang (2000) with Box-Muller normals.
 */
#include "../../include/sgpu_synth.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "../../include/sgpu.h"

typedef struct {
    uint64_t s[4];
} xrng;

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void xrng_seed(xrng *r, uint64_t seed, uint64_t stream)
{
    uint64_t x = seed ^ (stream * 0xD1342543DE82EF95ull + 0x632BE59BD9B4E019ull);
    for (int i = 0; i < 4; i++) {
        r->s[i] = splitmix64(&x);
    }
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t xnext(xrng *r)
{
    uint64_t *s = r->s;
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

static inline double xunif(xrng *r) { return (double)(xnext(r) >> 11) * 0x1p-53; }

static inline double xunif_open(xrng *r) { return ((double)(xnext(r) >> 11) + 0.5) * 0x1p-53; }

static double xnormal(xrng *r)
{
    double u1 = xunif_open(r), u2 = xunif_open(r);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static double xgamma(xrng *r, double a)
{
    if (a < 1.0) {
        double g = xgamma(r, a + 1.0);
        return g * exp(log(xunif_open(r)) / a);
    }
    double d = a - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
    for (;;) {
        double x = xnormal(r), v = 1.0 + c * x;
        if (v <= 0.0) {
            continue;
        }
        v = v * v * v;
        double u = xunif_open(r);
        if (u < 1.0 - 0.0331 * x * x * x * x || log(u) < 0.5 * x * x + d * (1.0 - v + log(v))) {
            return d * v;
        }
    }
}

static void xdirichlet(xrng *r, const double *a, int k, double *out)
{
    double s = 0.0;
    for (int i = 0; i < k; i++) {
        out[i] = xgamma(r, a[i]);
        s += out[i];
    }
    for (int i = 0; i < k; i++) {
        out[i] = s > 0 ? out[i] / s : 1.0 / k;
    }
}

static int nthreads_of(int32_t req)
{
    long n = req > 0 ? req : sysconf(_SC_NPROCESSORS_ONLN);
    if (n < 1) {
        n = 1;
    }
    if (n > 16) {
        n = 16;
    }
    return (int)n;
}

/* ---- SNP panel ----------------------------------------------------------- */

typedef struct {
    int32_t N, L, K;
    const double *pk;    /* [L][K] population frequencies of raw allele 1 */
    double qalpha, miss;
    uint64_t seed;
    uint32_t *packed;
    size_t lpad;
    double *qtrue;
    int32_t row0, row1;  /* word rows [row0, row1) */
} snp_job;

static void *snp_worker(void *arg)
{
    snp_job *j = (snp_job *)arg;
    double a[64], q[64];
    for (int k = 0; k < j->K; k++) {
        a[k] = j->qalpha;
    }
    for (int32_t g = j->row0; g < j->row1; g++) {
        uint32_t *row = j->packed + (size_t)g * j->lpad;
        for (int t = 0; t < 16; t++) {
            const int32_t i = g * 16 + t;
            if (i >= j->N) {
                break;
            }
            xrng r;
            xrng_seed(&r, j->seed, (uint64_t)i + 1);
            xdirichlet(&r, a, j->K, q);
            if (j->qtrue) {
                memcpy(j->qtrue + (size_t)i * j->K, q, sizeof(double) * j->K);
            }
            double cum[64];
            double s = 0.0;
            for (int k = 0; k < j->K; k++) {
                s += q[k];
                cum[k] = s;
            }
            for (int32_t l = 0; l < j->L; l++) {
                uint32_t code;
                if (xunif(&r) < j->miss) {
                    code = 3u;
                } else {
                    int nalt = 0;
                    for (int c = 0; c < 2; c++) {
                        const double u = xunif(&r) * s;
                        int z = 0;
                        while (z < j->K - 1 && u >= cum[z]) {
                            z++;
                        }
                        nalt += xunif(&r) < j->pk[(size_t)l * j->K + z];
                    }
                    code = (uint32_t)nalt;
                }
                row[l] = (row[l] & ~(3u << (2 * t))) | (code << (2 * t));
            }
        }
    }
    return NULL;
}

typedef struct {
    int32_t N, L, G;
    uint32_t *packed;
    size_t lpad;
    int32_t *nall;
    int8_t *geno;
    int32_t l0, l1;
} recode_job;

/* First-appearance recoding of a biallelic locus (datain.c:856-875): allele 0
 * is the allele of the first non-missing genotype (heterozygotes are listed
 * with the allele that becomes 0 first), so the locus flips when that
 * genotype is homozygous for raw allele 1. */
static void *snp_recode_worker(void *arg)
{
    recode_job *j = (recode_job *)arg;
    for (int32_t l = j->l0; l < j->l1; l++) {
        int first = -1, has0 = 0, has1 = 0;
        for (int32_t g = 0; g < j->G; g++) {
            const uint32_t w = j->packed[(size_t)g * j->lpad + l];
            for (int t = 0; t < 16 && g * 16 + t < j->N; t++) {
                const uint32_t code = (w >> (2 * t)) & 3u;
                if (code == 3u) {
                    continue;
                }
                if (first < 0) {
                    first = (int)code;
                }
                has0 |= code <= 1u;
                has1 |= code >= 1u;
            }
        }
        const int flip = first == 2;
        j->nall[l] = has0 + has1;
        for (int32_t g = 0; g < j->G; g++) {
            uint32_t *w = &j->packed[(size_t)g * j->lpad + l];
            for (int t = 0; t < 16 && g * 16 + t < j->N; t++) {
                uint32_t code = (*w >> (2 * t)) & 3u;
                if (code != 3u && flip) {
                    code = 2u - code;
                    *w = (*w & ~(3u << (2 * t))) | (code << (2 * t));
                }
                if (j->geno) {
                    const int32_t i = g * 16 + t;
                    int8_t *c0 = j->geno + (size_t)i * 2 * j->L;
                    int8_t *c1 = c0 + j->L;
                    if (code == 3u) {
                        c0[l] = c1[l] = -9;
                    } else {
                        c0[l] = code == 2u;
                        c1[l] = code >= 1u;
                    }
                }
            }
        }
    }
    return NULL;
}

int sgpu_synth_snp(int32_t N, int32_t L, int32_t Ktrue, double flo, double fhi, double qalpha,
                   double miss, uint64_t seed, int32_t nthreads, uint32_t *packed,
                   size_t packed_words, int8_t *geno, int32_t *nall, double *qtrue)
{
    if (N < 1 || L < 1 || Ktrue < 1 || Ktrue > 64 || !packed || !nall || !(flo > 0) ||
        !(fhi >= flo) || !(fhi < 1) || !(qalpha > 0) ||
        packed_words != sgpu_packed_words(SGPU_GENO_SNP2, N, L)) {
        return SGPU_ERR_ARG;
    }
    const size_t lpad = packed_words / (((size_t)N + 15) / 16);
    const int32_t G = (N + 15) / 16;
    double *pk = (double *)malloc(sizeof(double) * (size_t)L * Ktrue);
    if (!pk) {
        return SGPU_ERR_NOMEM;
    }
    xrng r;
    xrng_seed(&r, seed, 0);
    double F[64];
    for (int k = 0; k < Ktrue; k++) {
        F[k] = flo + (fhi - flo) * xunif(&r);
    }
    for (int32_t l = 0; l < L; l++) {
        const double p = 0.05 + 0.9 * xunif(&r);
        for (int k = 0; k < Ktrue; k++) {
            const double f = (1.0 - F[k]) / F[k];
            const double ga = xgamma(&r, p * f), gb = xgamma(&r, (1.0 - p) * f);
            const double s = ga + gb;
            pk[(size_t)l * Ktrue + k] = s > 0 ? ga / s : p;
        }
    }
    memset(packed, 0xff, packed_words * sizeof(uint32_t));
    const int nt = nthreads_of(nthreads);
    pthread_t th[16];
    snp_job jobs[16];
    const int32_t per = (G + nt - 1) / nt;
    for (int t = 0; t < nt; t++) {
        snp_job *j = &jobs[t];
        j->N = N; j->L = L; j->K = Ktrue; j->pk = pk; j->qalpha = qalpha; j->miss = miss;
        j->seed = seed; j->packed = packed; j->lpad = lpad; j->qtrue = qtrue;
        j->row0 = t * per < G ? t * per : G;
        j->row1 = (t + 1) * per < G ? (t + 1) * per : G;
        pthread_create(&th[t], NULL, snp_worker, j);
    }
    for (int t = 0; t < nt; t++) {
        pthread_join(th[t], NULL);
    }
    recode_job rj[16];
    const int32_t lper = (L + nt - 1) / nt;
    for (int t = 0; t < nt; t++) {
        recode_job *j = &rj[t];
        j->N = N; j->L = L; j->G = G; j->packed = packed; j->lpad = lpad; j->nall = nall;
        j->geno = geno;
        j->l0 = t * lper < L ? t * lper : L;
        j->l1 = (t + 1) * lper < L ? (t + 1) * lper : L;
        pthread_create(&th[t], NULL, snp_recode_worker, j);
    }
    for (int t = 0; t < nt; t++) {
        pthread_join(th[t], NULL);
    }
    free(pk);
    return SGPU_OK;
}

/* ---- microsatellite panel ------------------------------------------------ */

typedef struct {
    int32_t i0, i1, L, Ktrue, amax;
    double qalpha, miss;
    uint64_t seed;
    const int32_t *na;
    const double *cf;
    int8_t *geno;
    double *qtrue;
} ms_job;

typedef struct {
    int32_t N, L, l0, l1;
    int8_t *geno;
    int32_t *nall;
    int failed;
} ms_recode_job;

static void *ms_recode_worker(void *arg)
{
    ms_recode_job *j = (ms_recode_job *)arg;
    const int32_t n = j->l1 - j->l0;
    if (n <= 0) {
        return NULL;
    }
    int8_t *map = (int8_t *)malloc((size_t)n * 128);
    if (!map) {
        j->failed = 1;
        return NULL;
    }
    memset(map, -1, (size_t)n * 128);
    int32_t *nall = j->nall + j->l0;
    memset(nall, 0, sizeof(int32_t) * (size_t)n);
    for (int32_t i = 0; i < j->N; i++) {
        for (int c = 0; c < 2; c++) {
            int8_t *row = j->geno + ((size_t)i * 2 + c) * j->L + j->l0;
            for (int32_t l = 0; l < n; l++) {
                const int8_t x = row[l];
                if (x == -9) {
                    continue;
                }
                int8_t *m = map + (size_t)l * 128 + x;
                if (*m < 0) {
                    *m = (int8_t)nall[l]++;
                }
                row[l] = *m;
            }
        }
    }
    free(map);
    return NULL;
}

/* Individuals i0..i1-1 of the microsatellite panel; each has its own stream,
 * so the split over threads does not change the data. */
static void *ms_worker(void *arg)
{
    const ms_job *j = (const ms_job *)arg;
    const int32_t L = j->L, Ktrue = j->Ktrue, amax = j->amax;
    const int32_t *na = j->na;
    const double *cf = j->cf;
    const double miss = j->miss;
    const uint64_t seed = j->seed;
    int8_t *geno = j->geno;
    double *qtrue = j->qtrue;
    const int32_t ni = j->i1 - j->i0;
    if (ni <= 0) {
        return NULL;
    }
    /* every individual keeps its own stream; loci are visited in blocks whose
     * frequency table fits in cache, each individual still in locus order */
    xrng *st = (xrng *)malloc(sizeof(xrng) * (size_t)ni);
    double *cqs = (double *)malloc(sizeof(double) * (size_t)ni * Ktrue);
    double *sums = (double *)malloc(sizeof(double) * (size_t)ni);
    if (!st || !cqs || !sums) {
        free(st);
        free(cqs);
        free(sums);
        return (void *)1;
    }
    double qa[64], q[64];
    for (int k = 0; k < Ktrue; k++) {
        qa[k] = j->qalpha;
    }
    for (int32_t t = 0; t < ni; t++) {
        const int32_t i = j->i0 + t;
        xrng_seed(&st[t], seed, (uint64_t)i + 1);
        xdirichlet(&st[t], qa, Ktrue, q);
        if (qtrue) {
            memcpy(qtrue + (size_t)i * Ktrue, q, sizeof(double) * Ktrue);
        }
        double s = 0.0;
        for (int k = 0; k < Ktrue; k++) {
            s += q[k];
            cqs[(size_t)t * Ktrue + k] = s;
        }
        sums[t] = s;
    }
    const int32_t LB = 1024;
    for (int32_t lb = 0; lb < L; lb += LB) {
        const int32_t le = lb + LB < L ? lb + LB : L;
        for (int32_t t = 0; t < ni; t++) {
            const int32_t i = j->i0 + t;
            xrng *ri = &st[t];
            const double *cq = cqs + (size_t)t * Ktrue;
            const double s = sums[t];
            int8_t *c0 = geno + (size_t)i * 2 * L;
            int8_t *c1 = c0 + L;
            for (int32_t l = lb; l < le; l++) {
                if (xunif(ri) < miss) {
                    c0[l] = c1[l] = -9;
                    continue;
                }
                for (int c = 0; c < 2; c++) {
                    const double u = xunif(ri) * s;
                    int z = 0;
                    while (z < Ktrue - 1 && u >= cq[z]) {
                        z++;
                    }
                    const double *cz = cf + ((size_t)l * Ktrue + z) * amax;
                    const double v = xunif(ri) * cz[na[l] - 1];
                    int a = 0;
                    while (a < na[l] - 1 && v >= cz[a]) {
                        a++;
                    }
                    (c ? c1 : c0)[l] = (int8_t)a;
                }
            }
        }
    }
    free(st);
    free(cqs);
    free(sums);
    return NULL;
}


int sgpu_synth_microsat(int32_t N, int32_t L, int32_t amin, int32_t amax, int32_t Ktrue,
                        double qalpha, double miss, uint64_t seed, int8_t *geno, int32_t *nall,
                        double *qtrue)
{
    if (N < 1 || L < 1 || Ktrue < 1 || Ktrue > 64 || amin < 1 || amax < amin || amax > 126 ||
        !geno || !nall || !(qalpha > 0)) {
        return SGPU_ERR_ARG;
    }
    int32_t *na = (int32_t *)malloc(sizeof(int32_t) * (size_t)L);
    double *cf = (double *)malloc(sizeof(double) * (size_t)L * Ktrue * amax);
    if (!na || !cf) {
        free(na);
        free(cf);
        return SGPU_ERR_NOMEM;
    }
    xrng r;
    xrng_seed(&r, seed, 0);
    double ones[128], f[128];
    for (int a = 0; a < 128; a++) {
        ones[a] = 1.0;
    }
    for (int32_t l = 0; l < L; l++) {
        na[l] = amin + (int32_t)(xunif(&r) * (amax - amin + 1));
        for (int k = 0; k < Ktrue; k++) {
            xdirichlet(&r, ones, na[l], f);
            double s = 0.0;
            for (int a = 0; a < na[l]; a++) {
                s += f[a];
                cf[((size_t)l * Ktrue + k) * amax + a] = s;
            }
        }
    }
    int nt = (int)sysconf(_SC_NPROCESSORS_ONLN);
    nt = nt < 1 ? 1 : (nt > 16 ? 16 : nt);
    if (nt > N) {
        nt = N;
    }
    pthread_t th[16];
    ms_job jobs[16];
    for (int t = 0; t < nt; t++) {
        ms_job jb = {(int32_t)((int64_t)N * t / nt), (int32_t)((int64_t)N * (t + 1) / nt), L,
                     Ktrue, amax, qalpha, miss, seed, na, cf, geno, qtrue};
        jobs[t] = jb;
        pthread_create(&th[t], NULL, ms_worker, &jobs[t]);
    }
    int gen_failed = 0;
    for (int t = 0; t < nt; t++) {
        void *ret = NULL;
        pthread_join(th[t], &ret);
        gen_failed |= ret != NULL;
    }
    if (gen_failed) {
        free(na);
        free(cf);
        return SGPU_ERR_NOMEM;
    }
    /* first-appearance recoding over (ind, line) (datain.c:846-895): each
     * thread owns a range of loci and streams every row over it, with one
     * code table per locus */
    ms_recode_job rj[16];
    for (int t = 0; t < nt; t++) {
        ms_recode_job jb = {N, L, (int32_t)((int64_t)L * t / nt),
                            (int32_t)((int64_t)L * (t + 1) / nt), geno, nall, 0};
        rj[t] = jb;
        pthread_create(&th[t], NULL, ms_recode_worker, &rj[t]);
    }
    int failed = 0;
    for (int t = 0; t < nt; t++) {
        pthread_join(th[t], NULL);
        failed |= rj[t].failed;
    }
    if (failed) {
        free(na);
        free(cf);
        return SGPU_ERR_NOMEM;
    }
    int maxk = 0;
    for (int32_t l = 0; l < L; l++) {
        maxk = nall[l] > maxk ? nall[l] : maxk;
    }
    free(na);
    free(cf);
    return maxk > 0 ? maxk : 1;
}
