/* Result files of the `structure` command: <OUTFILE>_f, <OUTFILE>_q and the
 * INTERMEDSAVE files <OUTFILE>_<n>.
 *
 * The text is the reference's (OutPutResults, output.c:1354-1423, and the
 * printers it calls), so scripts that parse STRUCTURE output read these files
 * unchanged.  Only the admixture / F-model sections are produced, with the
 * ANCESTDIST probability intervals (no USEPOPINFO, LINKAGE or LOCPRIOR
 * columns: those models are not part of the sweep). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sgpu.h"
#include "../../include/structure_host.h"

#define SHOST_VERSION "2.3.4 (Jul 2012)"   /* structure.h:3, printed in the banner */

void shost_welcome(FILE *f)
{
    fprintf(f, "\n\n");
    fprintf(f, "----------------------------------------------------\n");
    fprintf(f, "STRUCTURE by Pritchard, Stephens and Donnelly (2000)\n");
    fprintf(f, "     and Falush, Stephens and Pritchard (2003)\n");
    fprintf(f, "       Code by Pritchard, Falush and Hubisz\n");
    fprintf(f, "             Version %s\n", SHOST_VERSION);
    fprintf(f, "----------------------------------------------------\n");
    fprintf(f, "\n\n");
    fflush(f);
}

/* (sum of squares - sum^2 / n) / (n - 1) (SampleVar, mymath.c:52-62). */
static double sample_var(double sumsq, double sum, long n)
{
    if (n == 1) {
        printf("Division by zero in variance calculation.\n");
        return 0;
    }
    return (sumsq - sum * sum / n) / (n - 1);
}

double shost_est_log_prob(double sumlikes, double sumsqlikes, int reps)
{
    const double mean = sumlikes / reps;
    return mean - sample_var(sumsqlikes, sumlikes, reps) / 2.0;
}

void shost_print_main_params(FILE *f, const shost_params *p, int rep, int argc, char **argv)
{
    fprintf(f, "\n");
    if (argc > 1) {
        fprintf(f, "Command line arguments:   ");
        for (int i = 0; i < argc; i++) {
            fprintf(f, "%s ", argv[i]);
        }
        fprintf(f, "\n");
    }
    fprintf(f, "Input File:    %s\n", p->datafile);
    if (f == stdout) {
        fprintf(f, "Output File:   %s_f\n", p->outfile);
    }
    fprintf(f, "\n");
    fprintf(f, "Run parameters:\n");
    fprintf(f, "   %d individuals\n", p->numinds);
    fprintf(f, "   %d loci\n", p->numloci);
    fprintf(f, "   %d populations assumed\n", p->maxpops);
    fprintf(f, "   %d Burn-in period\n", p->burnin);
    fprintf(f, "   %d Reps\n", rep - p->burnin);
    if (p->usepopinfo) {
        fprintf(f, "USEPOPINFO turned on\n");
        fprintf(f, "MIGRPRIOR = %1.4f\n", p->migrprior);
    }
    if (p->recessivealleles) {
        fprintf(f, "RECESSIVE ALLELES model used\n");
    }
    if (p->noadmix) {
        fprintf(f, "NO ADMIXTURE model assumed\n");
    }
    if (p->startatpopinfo) {
        fprintf(f, "STARTATPOPINFO turned on\n");
    }
    if (p->locprior) {
        fprintf(f, "LOCPRIOR model used\n");
    }
    if (!p->randomize) {
        fprintf(f, "RANDOMIZE turned off\n");
    }
    fprintf(f, "\n");
}

/* Row of K cluster sums normalised to proportions. */
static void q_totals(const shost_params *p, const shost_input *in, const double *qsum, int given,
                     int use_pop, double *sv, int *count)
{
    const int K = p->maxpops;
    for (int k = 0; k < K; k++) {
        sv[k] = 0.0;
    }
    *count = 0;
    for (int i = 0; i < in->N; i++) {
        if (use_pop && in->population[i] != given) {
            continue;
        }
        (*count)++;
        for (int k = 0; k < K; k++) {
            sv[k] += qsum[(size_t)i * K + k];
        }
    }
}

void shost_print_membership(FILE *f, const shost_params *p, const shost_input *in,
                            const double *qsum)
{
    const int K = p->maxpops;
    double *sv = (double *)calloc((size_t)K, sizeof(double));
    if (!sv) {
        fprintf(f, "Error assigning memory in function PrintMembership\n");
        return;
    }
    int n;
    if (p->popdata) {
        int lo = in->population[0], hi = in->population[0];
        for (int i = 1; i < in->N; i++) {
            lo = in->population[i] < lo ? in->population[i] : lo;
            hi = in->population[i] > hi ? in->population[i] : hi;
        }
        fprintf(f, "\n--------------------------------------------\n");
        fprintf(f, "Proportion of membership of each pre-defined\n");
        fprintf(f, " population in each of the %d clusters\n\n", K);
        fprintf(f, "Given    Inferred Clusters");
        for (int k = 3; k < K; k++) {
            fprintf(f, "       ");
        }
        fprintf(f, "       Number of\n");
        fprintf(f, " Pop    ");
        for (int k = 0; k < K; k++) {
            fprintf(f, "  %2d   ", k + 1);
        }
        for (int k = K; k < 3; k++) {
            fprintf(f, "       ");
        }
        fprintf(f, "   Individuals\n\n");
        for (int g = lo; g <= hi; g++) {
            q_totals(p, in, qsum, g, 1, sv, &n);
            double row = 0.0;
            for (int k = 0; k < K; k++) {
                row += sv[k];
            }
            if (row > 0.0) {
                fprintf(f, "%3d:     ", g);
                for (int k = 0; k < K; k++) {
                    fprintf(f, "%1.3f  ", sv[k] / row);
                }
                for (int k = K; k < 3; k++) {
                    fprintf(f, "       ");
                }
                fprintf(f, "    %3d\n", n);
            }
        }
    } else {
        q_totals(p, in, qsum, 0, 0, sv, &n);
        double row = 0.0;
        for (int k = 0; k < K; k++) {
            row += sv[k];
        }
        fprintf(f, "\n--------------------------------------------\n");
        fprintf(f, "Overall proportion of membership of the\n");
        fprintf(f, "sample in each of the %d clusters\n\n", K);
        fprintf(f, "Inferred Clusters\n");
        for (int k = 0; k < K; k++) {
            fprintf(f, " %2d    ", k + 1);
        }
        fprintf(f, "\n");
        for (int k = 0; k < K; k++) {
            fprintf(f, "%1.3f  ", sv[k] / row);
        }
        fprintf(f, "\n\n");
    }
    free(sv);
    fprintf(f, "--------------------------------------------\n");
}

/* Mean over loci of 0.5 (d11 + d22) - d12 (net) or 1 - d12 (gross), with
 * dij = sum_a Pi_a Pj_a of the posterior-mean frequencies (NETiv, GROSSiv,
 * output.c:470-519); psum holds sums over reps draws. */
static double distance(const shost_params *p, const shost_input *in, const double *psum, int k1,
                       int k2, int reps, int net)
{
    const int K = p->maxpops, A = in->maxalleles;
    double norm = (double)reps;
    norm *= norm;
    double sum = 0.0;
    for (int l = 0; l < in->L; l++) {
        const double *p1 = psum + ((size_t)l * K + k1) * A;
        const double *p2 = psum + ((size_t)l * K + k2) * A;
        double d1 = 0.0, d2 = 0.0, d12 = 0.0;
        for (int a = 0; a < in->nall[l]; a++) {
            if (net) {
                d1 += p1[a] * p1[a] / norm;
                d2 += p2[a] * p2[a] / norm;
            }
            d12 += p1[a] * p2[a] / norm;
        }
        sum += net ? 0.5 * (d1 + d2) - d12 : 1.0 - d12;
    }
    return sum / in->L;
}

void shost_print_net(FILE *f, const shost_params *p, const shost_input *in, const double *psum,
                     int reps, int format)
{
    const int K = p->maxpops;
    if (format == 0) {
        for (int i = 0; i < K - 1; i++) {
            for (int j = i + 1; j < K; j++) {
                fprintf(f, "%1.3f ", distance(p, in, psum, i, j, reps, 1));
            }
        }
        return;
    }
    fprintf(f, "\nAllele-freq. divergence among pops (Net nucleotide distance)");
    if (reps > 1) {
        fprintf(f, ",\ncomputed using point estimates of P");
    }
    fprintf(f, ".\n\n");
    fprintf(f, "     ");
    for (int j = 0; j < K; j++) {
        fprintf(f, "%2d      ", j + 1);
    }
    fprintf(f, "\n");
    for (int i = 0; i < K; i++) {
        fprintf(f, "%2d   ", i + 1);
        for (int j = 0; j < K; j++) {
            if (i == j) {
                fprintf(f, "   -    ");
            } else {
                fprintf(f, "%1.4f  ", distance(p, in, psum, i, j, reps, 1));
            }
        }
        fprintf(f, "\n");
    }
    fprintf(f, "\nAverage distances (expected heterozygosity) between individuals in same "
               "cluster:\n");
    for (int i = 0; i < K; i++) {
        fprintf(f, "cluster %2d  : %1.4f \n", i + 1, distance(p, in, psum, i, i, reps, 0));
    }
    fprintf(f, "\n");
}

void shost_print_sums(FILE *f, const shost_params *p, const shost_results *r)
{
    const int K = p->maxpops;
    const int n = r->rep - p->burnin;
    fprintf(f, "--------------------------------------------\n");
    if (p->computeprob) {
        if (n > 2) {
            fprintf(f, "Estimated Ln Prob of Data   = %1.1f\n",
                    shost_est_log_prob(r->sumlikes, r->sumsqlikes, n));
            fprintf(f, "Mean value of ln likelihood = %1.1f\n", r->sumlikes / n);
            fprintf(f, "Variance of ln likelihood   = %1.1f\n",
                    sample_var(r->sumsqlikes, r->sumlikes, n));
        } else {
            fprintf(f, "Mean value of ln likelihood = %1.1f\n", r->sumlikes / n);
        }
    }
    if (!p->noadmix && K > 1) {
        if (p->popalphas) {
            fprintf(f, "\n");
            for (int k = 0; k < K; k++) {
                fprintf(f, "Mean value of alpha_%d       = %1.4f\n", k + 1, r->alphasum[k] / n);
            }
        } else {
            fprintf(f, "Mean value of alpha         = %1.4f\n", r->alphasum[0] / n);
        }
    }
    if (p->inferlambda) {
        if (p->popspecificlambda) {
            for (int k = 0; k < K; k++) {
                fprintf(f, "\nMean value of lambda%d       = %1.4f\n", k + 1, r->lambdasum[k] / n);
            }
        } else {
            fprintf(f, "\nMean value of lambda        = %1.4f\n", r->lambdasum[0] / n);
        }
    }
    if (p->freqscorr) {
        if (p->onefst) {
            fprintf(f, "Mean value of Fst           = %1.4f\n", r->fstsum[0] / n);
        } else {
            fprintf(f, "\n");
            for (int k = 0; k < K; k++) {
                fprintf(f, "Mean value of Fst_%d         = %1.4f\n", k + 1, r->fstsum[k] / n);
            }
        }
    } else {
        fprintf(f, "Allele frequencies uncorrelated\n");
    }
    if (p->pfrompopflagonly) {
        fprintf(f, "Allele frequencies updated using individuals with POPFLAG=1 ONLY.\n");
    }
    fprintf(f, "\n");
}

static int missing_ind(const shost_input *in, int i)
{
    int n = 0;
    for (int c = 0; c < in->lines; c++) {
        const int32_t *g = in->geno + ((size_t)i * in->lines + c) * in->L;
        for (int l = 0; l < in->L; l++) {
            n += g[l] == in->missing;
        }
    }
    return n;
}

static int missing_loc(const shost_input *in, int l)
{
    int n = 0;
    for (int i = 0; i < in->N; i++) {
        for (int c = 0; c < in->lines; c++) {
            n += in->geno[((size_t)i * in->lines + c) * in->L + l] == in->missing;
        }
    }
    return n;
}

/* PrintAncestDist (output.c:762-797): per pop the box edges where the running
 * count of the histogram first exceeds (int)(n (1 - ANCESTPINT) / 2) and then
 * n (1 + ANCESTPINT) / 2, n the collected sweeps. */
static void print_ancest_dist(FILE *f, const shost_params *p, const int32_t *adist, int i, int n)
{
    const int K = p->maxpops, nb = p->numboxes;
    fprintf(f, "    ");
    for (int k = 0; k < K; k++) {
        const int32_t *h = adist + ((size_t)i * K + k) * nb;
        double sum = 0;
        int low = 0;
        for (int box = 0; box < nb; box++) {
            sum += h[box];
            if (!low && sum > (int)(n * (1.0 - p->ancestpint) / 2.0)) {
                fprintf(f, "(%1.3f,", (double)(box + 0.0) / nb);
                low = 1;
            }
            if (sum > (n * (1 + p->ancestpint) / 2.0)) {
                fprintf(f, "%1.3f) ", (double)(box + 1.0) / nb);
                break;
            }
        }
    }
}

void shost_print_q(FILE *f, const shost_params *p, const shost_input *in,
                   const shost_results *r)
{
    const int K = p->maxpops;
    const int n = r->rep - p->burnin;
    fprintf(f, "\n");
    fprintf(f, "Inferred ancestry of individuals:\n");
    fprintf(f, "    ");
    if (p->label) {
        fprintf(f, " %8s", "Label");
    }
    fprintf(f, " (%%Miss) ");
    if (p->popdata) {
        fprintf(f, "Pop");
    }
    fprintf(f, ":  ");
    fprintf(f, "Inferred clusters");
    if (p->ancestdist) {
        fprintf(f, " (and %d%c probability intervals)", (int)(100.0 * p->ancestpint), '%');
    }
    fprintf(f, "\n");
    for (int i = 0; i < in->N; i++) {
        const int miss = (int)(100 * missing_ind(in, i)) / (in->lines * in->L);
        fprintf(f, "%3d ", i + 1);
        if (p->label) {
            fprintf(f, "%8s ", in->label + (size_t)i * SHOST_LABELLEN);
        }
        if (miss < 10) {
            fprintf(f, " ");
        }
        fprintf(f, "  (%d) ", miss);
        fprintf(f, "  ");
        if (p->popdata) {
            fprintf(f, "%2d ", in->population[i]);
        }
        fprintf(f, ":  ");
        for (int k = 0; k < K; k++) {
            fprintf(f, "%1.3f ", r->qsum[(size_t)i * K + k] / n);
        }
        if (p->ancestdist && r->ancestdist) {
            print_ancest_dist(f, p, r->ancestdist, i, n);
        }
        fprintf(f, "\n");
    }
}

static void print_gene_name(FILE *f, const char *name)
{
    int i = 0;
    for (; i < SHOST_GENELEN && name[i] != '\0'; i++) {
        fputc(name[i], f);
    }
    if (i < SHOST_GENELEN) {
        if (i == 0) {
            fprintf(f, "XXX");
        }
        fprintf(f, " ");
    }
}

void shost_print_p(FILE *f, const shost_params *p, const shost_input *in,
                   const shost_results *r)
{
    const int K = p->maxpops, A = in->maxalleles;
    const int n = r->rep - p->burnin;
    fprintf(f, "\n\nEstimated Allele Frequencies in each cluster\n");
    if (p->freqscorr) {
        fprintf(f, "First column gives estimated ancestral frequencies\n");
    }
    fprintf(f, "\n\n");
    for (int l = 0; l < in->L; l++) {
        fprintf(f, "Locus %d : ", l + 1);
        if (p->markernames) {
            print_gene_name(f, in->markername + (size_t)l * SHOST_GENELEN);
        }
        fprintf(f, "\n");
        fprintf(f, "%d alleles\n", in->nall[l]);
        fprintf(f, "%2.1f%c missing data\n",
                (double)100.0 * missing_loc(in, l) / (in->lines * in->N), '%');
        for (int a = 0; a < in->nall[l]; a++) {
            fprintf(f, "%4d   ", in->translation[(size_t)l * A + a]);
            if (p->freqscorr) {
                fprintf(f, "(%1.3f) ", r->epssum[(size_t)l * A + a] / n);
            }
            for (int k = 0; k < K; k++) {
                fprintf(f, "%1.3f ", r->psum[((size_t)l * K + k) * A + a] / n);
            }
            fprintf(f, "\n");
        }
        fprintf(f, "\n");
    }
}

static void print_q_file(const shost_params *p, const shost_input *in, const shost_results *r)
{
    char name[SHOST_STRLEN + 20];
    snprintf(name, sizeof(name), "%s_q", p->outfile);
    FILE *f = fopen(name, "w");
    if (!f) {
        printf("WARNING: Unable to open output file %s.\n", name);
        return;
    }
    const int K = p->maxpops;
    const int n = r->rep - p->burnin;
    for (int i = 0; i < in->N; i++) {
        if (p->label) {
            fprintf(f, "%12s ", in->label + (size_t)i * SHOST_LABELLEN);
        } else {
            fprintf(f, "%4d ", i + 1);
        }
        if (p->popdata) {
            fprintf(f, "%2d ", in->population[i]);
        }
        for (int k = 0; k < K; k++) {
            fprintf(f, "%1.4f ", r->qsum[(size_t)i * K + k] / n);
        }
        fprintf(f, "\n");
    }
    fclose(f);
}

int shost_write_results(const shost_params *p, const shost_input *in, const shost_results *r,
                        int final, int argc, char **argv)
{
    char name[SHOST_STRLEN + 20];
    const int quiet = final == 2;
    final = final != 0;
    if (final) {
        snprintf(name, sizeof(name), "%s_f", p->outfile);
    } else {
        snprintf(name, sizeof(name), "%s_%d", p->outfile, r->saveindex);
    }
    FILE *f = fopen(name, "w");
    if (!f) {
        printf("WARNING: Unable to open output file %s.\n", name);
        return SGPU_ERR_ARG;
    }
    shost_welcome(f);
    if (final && !quiet) {
        printf("\nMCMC completed\n");
    }
    shost_print_main_params(f, p, r->rep, argc, argv);
    shost_print_membership(f, p, in, r->qsum);
    shost_print_net(f, p, in, r->psum, r->rep - p->burnin, 1);
    shost_print_sums(f, p, r);
    shost_print_q(f, p, in, r);
    shost_print_p(f, p, in, r);
    if (final && !quiet) {
        shost_print_q(stdout, p, in, r);
        if (p->printqhat) {
            print_q_file(p, in, r);
        }
        shost_print_main_params(stdout, p, r->rep, argc, argv);
        shost_print_sums(stdout, p, r);
        shost_print_membership(stdout, p, in, r->qsum);
    }
    shost_print_all_params(f, p);
    if (final && !quiet) {
        printf("Final results printed to file %s\n\n", name);
    }
    fclose(f);
    return SGPU_OK;
}

int32_t shost_save_freq(const shost_params *p)
{
    if (p->intermedsave <= 0) {
        return 0;
    }
    int32_t f = p->numreps / (p->intermedsave + 1);
    if (f * (p->intermedsave + 1) < p->numreps) {   /* one save too many otherwise */
        f++;
    }
    return f;
}

double shost_calc_like(const shost_input *in, int K, const double *Q, const double *P)
{
    const int N = in->N, L = in->L, A = in->maxalleles > 0 ? in->maxalleles : 1;
    const double clamp = 1e-50;   /* sqrt(UNDERFLO), CalcLike.c:99 */
    double loglike = 0.0;
    for (int i = 0; i < N; i++) {
        double run = 1.0, ind = 0.0;
        for (int c = 0; c < in->lines && c < 2; c++) {
            for (int l = 0; l < L; l++) {
                const int x = in->geno[((size_t)i * in->lines + c) * L + l];
                if (x == in->missing || x < 0 || x >= in->nall[l]) {
                    continue;
                }
                double term = 0.0;
                for (int k = 0; k < K; k++) {
                    term += Q[(size_t)i * K + k] * P[((size_t)l * K + k) * A + x];
                }
                run *= term > clamp ? term : clamp;
                if (run < clamp) {
                    ind += log(run);
                    run = 1.0;
                }
            }
        }
        loglike += ind + log(run);
    }
    return loglike;
}
