/* `structure`: the reference's command-line program (structure.c:148-742)
 * with the admixture sweep on the GPU through sgpu.h.
 *
 *   structure [-m mainparams] [-e extraparams] [-i infile] [-o outfile]
 *             [-K k] [-L loci] [-N inds] [-D seed]
 *             [-B kmin kmax replicates]
 *
 * reads the same parameter and data files and writes the same <OUTFILE>_f
 * (and, with PRINTQHAT, <OUTFILE>_q).  -B runs the K-sweep users script
 * around structure (K = kmin..kmax, `replicates` seeds each) as one batch of
 * chains on the GPU, writes <OUTFILE>_K<k>_R<r>_f per chain, and aligns the
 * replicates of each K into <OUTFILE>_K<k>_summary. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/sgpu.h"
#include "../../include/structure_align.h"
#include "../../include/structure_host.h"

static int fail_sgpu(int rc)
{
    printf("sgpu error %d: %s\n", rc, sgpu_last_error());
    printf("\nExiting the program due to error(s) listed above.\n\n");
    return 1;
}

static int unsupported(const shost_params *p)
{
    const struct {
        int on;
        const char *what;
    } u[] = {{p->linkage, "LINKAGE"}, {p->usepopinfo, "USEPOPINFO"}, {p->locprior, "LOCPRIOR"},
             {p->noadmix, "NOADMIX"}, {p->recessivealleles, "RECESSIVEALLELES"},
             {p->ploidy != 2, "PLOIDY other than 2"}};
    int bad = 0;
    for (size_t i = 0; i < sizeof(u) / sizeof(u[0]); i++) {
        if (u[i].on) {
            printf("%s is not supported by the GPU admixture sweep\n", u[i].what);
            bad = 1;
        }
    }
    if (p->maxpops < 1) {
        printf("Need to set MAXPOPS>0 (mainparams or -K)\n");
        bad = 1;
    }
    if (p->pfrompopflagonly && !p->popflag) {   /* CheckParamCombinations, Util.c:75-78 */
        printf("PFROMPOPFLAGONLY can only be turned on when the data file contains POPFLAG "
               "data\n");
        bad = 1;
    }
    if (p->ancestdist && p->numboxes < 1) {
        printf("NUMBOXES must be at least 1 with ANCESTDIST\n");
        bad = 1;
    }
    return bad;
}

static sgpu_model model_of(const shost_params *p, int K, int A, uint32_t seed, uint32_t chain)
{
    sgpu_model m;
    memset(&m, 0, sizeof(m));
    m.N = p->numinds;
    m.L = p->numloci;
    m.K = K;
    m.A = A;
    m.burnin = p->burnin;
    m.numreps = p->numreps;
    m.freqscorr = p->freqscorr;
    m.onefst = p->onefst;
    m.inferalpha = p->inferalpha;
    m.popalphas = p->popalphas;
    m.unifprioralpha = p->unifprioralpha;
    m.inferlambda = p->inferlambda;
    m.popspecificlambda = p->popspecificlambda;
    m.metrofreq = p->metrofreq;
    m.computeprob = p->computeprob;
    m.alpha0 = p->alpha;
    m.lambda0 = p->lambda;
    m.alphapropsd = p->alphapropsd;
    m.alphamax = p->alphamax;
    m.alphapriora = p->alphapriora;
    m.alphapriorb = p->alphapriorb;
    m.fpriormean = p->fpriormean;
    m.fpriorsd = p->fpriorsd;
    m.seed = seed;
    m.chain = chain;
    m.ancestdist = p->ancestdist ? 1 : 0;
    m.numboxes = p->numboxes;
    m.pfrompopflagonly = p->pfrompopflagonly ? 1 : 0;
    return m;
}

/* Packed genotypes: 2-bit when every site is representable, else bytes
 * straight from the recoded int32 genotypes (up to 255 alleles per locus). */
static int pack(const shost_input *in, sgpu_data *d, uint32_t **words, int *A)
{
    int rc = SGPU_OK;
    *A = in->maxalleles > 0 ? in->maxalleles : 1;
    int fmt = SGPU_GENO_BYTE;
    size_t nw = 0;
    *words = NULL;
    if (*A == 2) {
        const size_t n = (size_t)in->N * 2 * in->L;
        int8_t *g8 = (int8_t *)malloc(n);
        if (!g8) {
            return SGPU_ERR_NOMEM;
        }
        nw = sgpu_packed_words(SGPU_GENO_SNP2, in->N, in->L);
        *words = (uint32_t *)malloc(nw * sizeof(uint32_t));
        if (*words && shost_geno_int8(in, g8) == SGPU_OK &&
            sgpu_pack_snp2(g8, in->N, in->L, *words, nw) == SGPU_OK) {
            fmt = SGPU_GENO_SNP2;
        }
        free(g8);
    }
    if (fmt == SGPU_GENO_BYTE) {
        free(*words);
        if (in->lines != 2) {
            return SGPU_ERR_UNSUPPORTED;
        }
        nw = sgpu_packed_words(SGPU_GENO_BYTE, in->N, in->L);
        *words = (uint32_t *)malloc(nw * sizeof(uint32_t));
        rc = *words ? sgpu_pack_byte32(in->geno, in->N, in->L, in->missing, *words, nw)
                    : SGPU_ERR_NOMEM;
        if (rc == SGPU_ERR_ARG) {
            printf("allele codes do not fit the byte genotype format (at most 255 alleles)\n");
        }
    }
    d->format = fmt;
    d->geno = *words;
    d->geno_words = nw;
    d->nall = in->nall;
    d->popflag = in->popflag;   /* read by PFROMPOPFLAGONLY (POPFLAG=1 files only) */
    printf("Genotypes packed as %s (%zu words)\n", fmt == SGPU_GENO_SNP2 ? "2-bit SNP" : "bytes",
           nw);
    return rc;
}

/* Sums of chain c of ctx after the run, in host buffers. */
typedef struct {
    double *qsum, *psum, *epssum, *fstsum, *alphasum, *lambdasum, like[3];
    double *alpha, *fst;
    int32_t *adist;   /* ANCESTDIST histogram, or NULL */
} sums_t;

static void free_sums(sums_t *s)
{
    free(s->adist);
    s->adist = NULL;
    free(s->qsum);
    free(s->psum);
    free(s->epssum);
    free(s->fstsum);
    free(s->alphasum);
    free(s->lambdasum);
    free(s->alpha);
    free(s->fst);
}

static int read_sums(sgpu_ctx *ctx, int c, const sgpu_model *m, sums_t *s)
{
    const size_t N = m->N, L = m->L, K = m->K, A = m->A;
    memset(s, 0, sizeof(*s));
    s->qsum = (double *)malloc(sizeof(double) * N * K);
    s->psum = (double *)malloc(sizeof(double) * L * K * A);
    s->epssum = (double *)malloc(sizeof(double) * L * A);
    s->fstsum = (double *)malloc(sizeof(double) * K);
    s->alphasum = (double *)malloc(sizeof(double) * K);
    s->lambdasum = (double *)malloc(sizeof(double) * K);
    s->alpha = (double *)malloc(sizeof(double) * K);
    s->fst = (double *)malloc(sizeof(double) * K);
    if (!s->qsum || !s->psum || !s->epssum || !s->fstsum || !s->alphasum || !s->lambdasum ||
        !s->alpha || !s->fst) {
        return SGPU_ERR_NOMEM;
    }
    int rc;
    if ((rc = sgpu_get_chain_state(ctx, c, SGPU_QSUM, s->qsum, sizeof(double) * N * K)) ||
        (rc = sgpu_get_chain_state(ctx, c, SGPU_PSUM, s->psum, sizeof(double) * L * K * A)) ||
        (rc = sgpu_get_chain_state(ctx, c, SGPU_EPSSUM, s->epssum, sizeof(double) * L * A)) ||
        (rc = sgpu_get_chain_state(ctx, c, SGPU_FSTSUM, s->fstsum, sizeof(double) * K)) ||
        (rc = sgpu_get_chain_state(ctx, c, SGPU_ALPHASUM, s->alphasum, sizeof(double) * K)) ||
        (rc = sgpu_get_chain_state(ctx, c, SGPU_LAMBDASUM, s->lambdasum, sizeof(double) * K)) ||
        (rc = sgpu_get_chain_state(ctx, c, SGPU_ALPHA, s->alpha, sizeof(double) * K)) ||
        (rc = sgpu_get_chain_state(ctx, c, SGPU_FST, s->fst, sizeof(double) * K)) ||
        (rc = sgpu_get_chain_state(ctx, c, SGPU_LIKE, s->like, sizeof(s->like)))) {
        return rc;
    }
    if (m->ancestdist) {
        const size_t nb = N * K * (size_t)m->numboxes;
        s->adist = (int32_t *)malloc(sizeof(int32_t) * nb);
        if (!s->adist) {
            return SGPU_ERR_NOMEM;
        }
        if ((rc = sgpu_get_chain_state(ctx, c, SGPU_ANCESTDIST, s->adist, sizeof(int32_t) * nb))) {
            return rc;
        }
    }
    return SGPU_OK;
}

static shost_results results_of(int rep, const sums_t *s)
{
    shost_results r;
    memset(&r, 0, sizeof(r));
    r.rep = rep;
    r.qsum = s->qsum;
    r.psum = s->psum;
    r.epssum = s->epssum;
    r.fstsum = s->fstsum;
    r.alphasum = s->alphasum;
    r.lambdasum = s->lambdasum;
    r.sumlikes = s->like[1];
    r.sumsqlikes = s->like[2];
    r.ancestdist = s->adist;
    return r;
}

/* PRINTLIKES (structure.c:663-667, PrintLike output.c:187-200): the lnP(D) of
 * the sweep just run; a collecting sweep has it from the device, a burn-in
 * sweep does not accumulate it and it is computed on the host from Q and P. */
static int print_like(sgpu_ctx *ctx, const shost_params *p, const shost_input *in,
                      const sgpu_model *m, int rep)
{
    int rc;
    if (rep + 1 > p->burnin && p->computeprob) {
        double like[3];
        if ((rc = sgpu_get_state(ctx, SGPU_LIKE, like, sizeof(like)))) {
            return rc;
        }
        printf("%6.0f\n", like[0]);
        return SGPU_OK;
    }
    const size_t nq = (size_t)m->N * m->K, np = (size_t)m->L * m->K * m->A;
    double *Q = (double *)malloc(sizeof(double) * nq), *P = (double *)malloc(sizeof(double) * np);
    rc = (Q && P) ? SGPU_OK : SGPU_ERR_NOMEM;
    if (rc == SGPU_OK && (rc = sgpu_get_state(ctx, SGPU_Q, Q, sizeof(double) * nq)) == SGPU_OK &&
        (rc = sgpu_get_state(ctx, SGPU_P, P, sizeof(double) * np)) == SGPU_OK) {
        printf("%6.0f\n", shost_calc_like(in, m->K, Q, P));
    }
    free(Q);
    free(P);
    return rc;
}

/* One progress line (the columns of PrintUpdate, output.c:218-375). */
static void progress(const shost_params *p, int rep, const sums_t *s)
{
    printf("%5d:    ", rep);
    if (p->popalphas) {
        for (int k = 0; k < p->maxpops; k++) {
            printf("%1.3f  ", s->alpha[k]);
        }
    } else {
        printf("%1.3f  ", s->alpha[0]);
    }
    if (p->freqscorr) {
        printf("  ");
        const int nf = p->onefst ? 1 : p->maxpops;
        for (int k = 0; k < nf; k++) {
            printf("%1.3f ", s->fst[k]);
        }
        printf("   ");
    } else {
        printf("  ");
    }
    if (p->computeprob) {
        const int n = rep - p->burnin;
        if (n > 2) {
            printf("  %.0f  ", s->like[0]);
            printf("  %.0f ", shost_est_log_prob(s->like[1], s->like[2], n));
        } else {
            printf("  --  ");
        }
    }
    printf("\n");
    fflush(stdout);
}

static int run_one(shost_params *p, shost_input *in, const sgpu_data *d, int A, uint32_t seed,
                   int argc, char **argv)
{
    /* Philox key (SEED, K): the same chain as replicate SEED of a -B batch */
    sgpu_model m = model_of(p, p->maxpops, A, seed, (uint32_t)p->maxpops);
    sgpu_ctx *ctx = NULL;
    int rc;
    if ((rc = sgpu_create(&m, d, &ctx)) || (rc = sgpu_init_state(ctx))) {
        sgpu_destroy(ctx);
        return fail_sgpu(rc);
    }
    printf("\n\n--------------------------------------\n\n");
    printf("Finished initialization; starting MCMC \n");
    printf("%d iterations + %d burnin\n\n", p->numreps, p->burnin);
    const int total = p->numreps + p->burnin;
    int rep = 0;
    sums_t s;
    memset(&s, 0, sizeof(s));
    s.alpha = (double *)calloc((size_t)m.K, sizeof(double));
    s.fst = (double *)calloc((size_t)m.K, sizeof(double));
    if (!s.alpha || !s.fst) {
        rc = SGPU_ERR_NOMEM;
    }
    /* sweeps run in stretches that end where the reference reads the device: a
     * progress line (UPDATEFREQ), an INTERMEDSAVE file, or every sweep with
     * PRINTLIKES */
    const int savefreq = shost_save_freq(p);
    while (rc == SGPU_OK && rep < total) {
        int n = p->updatefreq - rep % p->updatefreq;
        if (p->printlikes) {
            n = 1;
        }
        if (savefreq > 0) {
            /* next sweep r (0-based) with (r + 1 - BURNIN) % savefreq == 0 */
            int r = rep < p->burnin ? p->burnin : rep;
            const int into = (r + 1 - p->burnin) % savefreq;
            r += into ? savefreq - into : 0;
            if (r + 1 - rep < n) {
                n = r + 1 - rep;
            }
        }
        if (rep + n > total) {
            n = total - rep;
        }
        if ((rc = sgpu_run_sweeps(ctx, (uint32_t)rep, (uint32_t)n))) {
            break;
        }
        rep += n;
        const int last = rep - 1;   /* the sweep just run */
        if (savefreq && last + 1 > p->burnin && (last + 1 - p->burnin) % savefreq == 0 &&
            last + 1 != total) {
            /* an intermediate result file (structure.c:640-661) */
            sums_t t;
            if ((rc = read_sums(ctx, 0, &m, &t)) == SGPU_OK) {
                shost_results r = results_of(rep, &t);
                r.saveindex = (rep - p->burnin) / savefreq;
                rc = shost_write_results(p, in, &r, 0, argc, argv);
            }
            free_sums(&t);
            if (rc) {
                break;
            }
        }
        if (p->printlikes && (rc = print_like(ctx, p, in, &m, last))) {
            break;
        }
        if (rep % p->updatefreq != 0 && rep != total) {
            continue;
        }
        /* the progress line needs alpha, F and the likelihood sums only */
        const size_t K = (size_t)m.K;
        if ((rc = sgpu_get_state(ctx, SGPU_ALPHA, s.alpha, sizeof(double) * K)) ||
            (rc = sgpu_get_state(ctx, SGPU_FST, s.fst, sizeof(double) * K)) ||
            (rc = sgpu_get_state(ctx, SGPU_LIKE, s.like, sizeof(s.like)))) {
            break;
        }
        if (rep == p->burnin) {
            printf("\nBURNIN completed\n");
        }
        progress(p, rep, &s);
    }
    if (rc == SGPU_OK) {
        free_sums(&s);
        rc = read_sums(ctx, 0, &m, &s);
    }
    if (rc == SGPU_OK) {
        shost_results r = results_of(rep, &s);
        rc = shost_write_results(p, in, &r, 1, argc, argv);
    }
    free_sums(&s);
    sgpu_destroy(ctx);
    return rc ? fail_sgpu(rc) : 0;
}

static int run_ksweep(shost_params *p, shost_input *in, const sgpu_data *d, int A, uint32_t seed,
                      int kmin, int kmax, int reps, int argc, char **argv)
{
    if (kmin < 1 || kmax < kmin || reps < 1) {
        printf("-B needs 1 <= kmin <= kmax and replicates >= 1\n");
        return 1;
    }
    const int nk = kmax - kmin + 1, nc = nk * reps;
    sgpu_model *models = (sgpu_model *)calloc((size_t)nc, sizeof(sgpu_model));
    int32_t *doc = (int32_t *)calloc((size_t)nc, sizeof(int32_t));
    if (!models || !doc) {
        free(models);
        free(doc);
        return fail_sgpu(SGPU_ERR_NOMEM);
    }
    for (int k = 0; k < nk; k++) {
        for (int r = 0; r < reps; r++) {
            models[k * reps + r] = model_of(p, kmin + k, A, seed + (uint32_t)r, (uint32_t)(kmin + k));
        }
    }
    sgpu_ctx *ctx = NULL;
    int rc;
    if ((rc = sgpu_create_batch(models, nc, d, 1, doc, NULL, &ctx)) ||
        (rc = sgpu_init_state(ctx))) {
        sgpu_destroy(ctx);
        free(models);
        free(doc);
        return fail_sgpu(rc);
    }
    printf("\n%d chains (K = %d..%d, %d replicates each) in one batch\n", nc, kmin, kmax, reps);
    printf("%d iterations + %d burnin\n\n", p->numreps, p->burnin);
    const int total = p->numreps + p->burnin;
    const clock_t t0 = clock();
    rc = sgpu_run_sweeps(ctx, 0, (uint32_t)total);
    if (rc == SGPU_OK) {
        rc = sgpu_synchronize(ctx);
    }
    printf("batch done in %.1f s of host time\n", (double)(clock() - t0) / CLOCKS_PER_SEC);
    const char *base = p->outfile;
    char saved[SHOST_STRLEN + 1];
    snprintf(saved, sizeof(saved), "%s", base);
    for (int k = 0; k < nk && rc == SGPU_OK; k++) {
        const int K = kmin + k;
        const size_t nq = (size_t)in->N * K;
        double *qbars = (double *)malloc(sizeof(double) * nq * reps);
        double *lnp = (double *)malloc(sizeof(double) * reps);
        if (!qbars || !lnp) {
            free(qbars);
            free(lnp);
            rc = SGPU_ERR_NOMEM;
            break;
        }
        shost_params pk = *p;
        pk.maxpops = K;
        for (int r = 0; r < reps && rc == SGPU_OK; r++) {
            sums_t s;
            if ((rc = read_sums(ctx, k * reps + r, &models[k * reps + r], &s)) == SGPU_OK) {
                char tag[32];
                snprintf(tag, sizeof(tag), "_K%d_R%d", K, r + 1);
                if (strlen(saved) + strlen(tag) >= sizeof(pk.outfile)) {
                    printf("output file name too long\n");
                    rc = SGPU_ERR_ARG;
                    free_sums(&s);
                    break;
                }
                strcpy(pk.outfile, saved);
                strcat(pk.outfile, tag);
                pk.seed = (int32_t)(seed + (uint32_t)r);
                shost_results res = results_of(total, &s);
                rc = shost_write_results(&pk, in, &res, 2 /* <name>_f, no stdout */, argc, argv);
                const int n = total - p->burnin;
                for (size_t j = 0; j < nq; j++) {
                    qbars[(size_t)r * nq + j] = s.qsum[j] / n;
                }
                lnp[r] = n > 2 ? shost_est_log_prob(s.like[1], s.like[2], n) : s.like[1] / n;
            }
            free_sums(&s);
        }
        if (rc == SGPU_OK) {
            double *qmean = (double *)malloc(sizeof(double) * nq);
            double *qsd = (double *)malloc(sizeof(double) * nq);
            int32_t *perms = (int32_t *)malloc(sizeof(int32_t) * (size_t)reps * K);
            double m = 0, se = 0;
            if (!qmean || !qsd || !perms ||
                (rc = shost_replicate_summary(qbars, lnp, reps, in->N, K, qmean, qsd, &m, &se,
                                              perms))) {
                rc = rc ? rc : SGPU_ERR_NOMEM;
            } else {
                char name[SHOST_STRLEN + 40];
                snprintf(name, sizeof(name), "%s_K%d_summary", saved, K);
                FILE *f = fopen(name, "w");
                if (f) {
                    fprintf(f, "K = %d, %d replicates aligned to replicate 1\n", K, reps);
                    fprintf(f, "Estimated Ln Prob of Data: mean %.1f, Monte Carlo s.e. %.1f\n", m, se);
                    for (int r = 0; r < reps; r++) {
                        fprintf(f, "replicate %d: Ln P(D) = %.1f, clusters", r + 1, lnp[r]);
                        for (int a = 0; a < K; a++) {
                            fprintf(f, " %d", perms[(size_t)r * K + a] + 1);
                        }
                        fprintf(f, "\n");
                    }
                    fprintf(f, "\nAligned mean ancestry (sd over replicates):\n");
                    for (int i = 0; i < in->N; i++) {
                        fprintf(f, "%4d ", i + 1);
                        for (int a = 0; a < K; a++) {
                            fprintf(f, "%1.4f (%1.4f) ", qmean[(size_t)i * K + a],
                                    qsd[(size_t)i * K + a]);
                        }
                        fprintf(f, "\n");
                    }
                    fclose(f);
                }
                printf("K=%2d: Ln P(D) %.1f +- %.1f over %d replicates -> %s\n", K, m, se, reps,
                       name);
            }
            free(qmean);
            free(qsd);
            free(perms);
        }
        free(qbars);
        free(lnp);
    }
    sgpu_destroy(ctx);
    free(models);
    free(doc);
    return rc ? fail_sgpu(rc) : 0;
}

int main(int argc, char **argv)
{
    shost_welcome(stdout);
    shost_params p;
    if (shost_get_params(&p, argc, argv) || unsupported(&p)) {
        printf("\nExiting the program due to error(s) listed above.\n\n");
        return 1;
    }
    int kmin = 0, kmax = 0, reps = 0;
    for (int i = 0; i + 3 < argc; i++) {
        if (strcmp(argv[i], "-B") == 0) {
            kmin = atoi(argv[i + 1]);
            kmax = atoi(argv[i + 2]);
            reps = atoi(argv[i + 3]);
        }
    }
    /* Randomize (ran.c:113-125): a time seed unless RANDOMIZE is off and SEED
     * is not the shipped placeholder 2245; the seed is appended to seed.txt */
    if (p.randomize || p.seed == 2245) {
        p.seed = (int32_t)time(NULL);
    }
    FILE *sf = fopen("seed.txt", "a");
    if (sf) {
        fprintf(sf, "%i\n", p.seed);
        fclose(sf);
    }
    shost_input in;
    if (shost_read_input(&p, &in)) {
        printf("\nExiting the program due to error(s) listed above.\n\n");
        return 1;
    }
    sgpu_data d;
    uint32_t *words = NULL;
    int A = 0;
    int rc = pack(&in, &d, &words, &A);
    if (rc) {
        shost_free_input(&in);
        free(words);
        return fail_sgpu(rc);
    }
    int status = reps > 0 ? run_ksweep(&p, &in, &d, A, (uint32_t)p.seed, kmin, kmax, reps, argc, argv)
                          : run_one(&p, &in, &d, A, (uint32_t)p.seed, argc, argv);
    free(words);
    shost_free_input(&in);
    printf("Structure seed: %d\n", p.seed);
    return status;
}
