/*
 * structure_host.h -- host side of the `structure` command built on sgpu.h:
 * the mainparams/extraparams reader, the genotype file reader with
 * first-appearance allele recoding, and the result-file writer.
 *
 * Each function names the reference code whose behaviour it keeps
 * (pmelsted/structure src/params.c, datain.c, output.c).  Plain C; errors
 * are returned as SGPU_ERR_* codes after printing the reference's message to
 * stdout (the reference prints the same text and exits through Kill(),
 * Util.c:29-33).
 */
#ifndef STRUCTURE_HOST_H
#define STRUCTURE_HOST_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHOST_STRLEN 200   /* file names (structure.h:9) */
#define SHOST_GENELEN 15   /* marker names (structure.h:10) */
#define SHOST_LABELLEN 12  /* individual labels (structure.h:11) */
#define SHOST_ALLELLEN 10  /* integer tokens of the data file (structure.h:13) */

/* Every parameter key of mainparams/extraparams (structure.h:70-131), with
 * the reference's defaults (PresetValues, params.c:571-700). */
typedef struct {
    char datafile[SHOST_STRLEN + 1], outfile[SHOST_STRLEN + 1];
    int32_t numinds, numloci, missing, label, popdata, popflag, phenotype, extracols;
    int32_t markernames, phaseinfo, mapdistances, markovphase, ploidy, onerowperind;
    int32_t recessivealleles, notambiguous, phased, locdata;
    int32_t maxpops, burnin, numreps;
    int32_t usepopinfo, pfrompopflagonly, inferalpha, inferlambda, popspecificlambda, popalphas;
    int32_t computeprob, noadmix, admburnin, printqhat, printqsum, updatefreq, printlikes;
    int32_t intermedsave, printkld, printnet, printlambda, echodata, ancestdist, numboxes;
    int32_t gensback, freqscorr, onefst, unifprioralpha, startatpopinfo, randomize, linkage;
    int32_t metrofreq, reporthitrate;
    double ancestpint, migrprior, alpha, log10rpropsd, log10rmin, log10rmax, log10rstart;
    double fpriormean, fpriorsd, lambda, alphamax, alphapriora, alphapriorb, alphapropsd;
    int32_t noqs, posterior, picturefile, individualr;
    int32_t numsimstats, poolfreq, locusxonly, phenotypecol, missingpheno;
    double emerror;
    int32_t locprior, updatelocprior, locispop, printlocprior;
    double locpriorinit, maxlocprior, locpriorstep;
    int32_t seed;
    int32_t origmissing;   /* MISSING as given, before recoding (ORIGMISSING) */
} shost_params;

/* PresetValues (params.c:571-700). */
void shost_preset(shost_params *p);

/* ReadFile / GetParamValue / SetValue (params.c:225-570): every "#define KEY
 * value" in the file; unknown keys are reported and ignored. */
int shost_read_params(shost_params *p, const char *path);

/* InputFileNames (params.c:147-215): the file named after flag ("-m", "-e")
 * or the default name; writes at most n bytes to out. */
void shost_param_file_name(int argc, char **argv, const char *flag, char *out, size_t n);

/* CommandLineValues (params.c:84-127): -i -o -K -L -N -r -c -D. */
int shost_command_line(shost_params *p, int argc, char **argv);

/* CheckIfMissing (params.c:703-768). */
int shost_check(const shost_params *p);

/* SetUpdateFreq (params.c:770-795). */
int32_t shost_update_freq(const shost_params *p);

/* GetParams (params.c:49-82) without STRAT: preset, mainparams, extraparams,
 * command line, checks, UPDATEFREQ. */
int shost_get_params(shost_params *p, int argc, char **argv);

/* PrintAllParams (params.c:797-857). */
void shost_print_all_params(FILE *f, const shost_params *p);

/* The genotype file after ReadInputFile (datain.c:32-50, ReadData :62-333) and
 * CountAlleles (datain.c:824-904). */
typedef struct {
    int32_t N, L, lines;
    int32_t maxalleles;        /* FindMaxAlleles (datain.c:768-820) */
    int32_t missing;           /* missing code after recoding (-9 unless outside) */
    int32_t *geno;             /* [N][lines][L] recoded alleles 0..k-1 or missing */
    int32_t *nall;             /* [L] NumAlleles */
    int32_t *translation;      /* [L][maxalleles] original allele of each code */
    char *label;               /* [N][SHOST_LABELLEN] */
    int32_t *population, *popflag, *phenotype, *location;   /* [N] */
    char *markername;          /* [L][SHOST_GENELEN] */
    double *mapdistance;       /* [L] */
} shost_input;

/* Read p->datafile and recode.  Prints the reference's warnings.  Supported:
 * PLOIDY 2, no RECESSIVEALLELES, no PHASEINFO rows beyond skipping. */
int shost_read_input(shost_params *p, shost_input *in);
void shost_free_input(shost_input *in);

/* Recoded genotypes as int8 [N][2][L] with -9 for missing, the input of
 * sgpu_pack_snp2 / sgpu_pack_byte.  out holds N*2*L bytes. */
int shost_geno_int8(const shost_input *in, int8_t *out);

/* Posterior sums of a run (device state after the last sweep so far). */
typedef struct {
    int32_t rep;               /* BURNIN + number of collected sweeps */
    const double *qsum;        /* [N][K] */
    const double *psum;        /* [L][K][A] with A = in->maxalleles */
    const double *epssum;      /* [L][A] (F model) */
    const double *fstsum, *alphasum, *lambdasum;   /* [K] */
    double sumlikes, sumsqlikes;
    const int32_t *ancestdist; /* [N][K][NUMBOXES] with ANCESTDIST, else NULL */
    int32_t saveindex;         /* final = 0: (rep - BURNIN) / savefreq, the file's number */
} shost_results;

/* OutPutResults (output.c:1354-1423): final = 1 writes <OUTFILE>_f, the
 * stdout summary and, with PRINTQHAT, the <OUTFILE>_q file (PrintQFile,
 * output.c:915-953); final = 2 writes <OUTFILE>_f only; final = 0 writes the
 * intermediate file <OUTFILE>_<saveindex> of INTERMEDSAVE (structure.c:640-661). */
int shost_write_results(const shost_params *p, const shost_input *in, const shost_results *r,
                        int final, int argc, char **argv);

/* The parts of the result file, for callers that print them elsewhere. */
void shost_welcome(FILE *f);                                       /* Util.c:15-27 */
void shost_print_main_params(FILE *f, const shost_params *p, int rep, int argc, char **argv);
void shost_print_membership(FILE *f, const shost_params *p, const shost_input *in,
                            const double *qsum);                  /* output.c:1213-1336 */
void shost_print_net(FILE *f, const shost_params *p, const shost_input *in, const double *psum,
                     int reps, int format);                        /* output.c:543-596 */
void shost_print_sums(FILE *f, const shost_params *p, const shost_results *r); /* :1022-1144 */
void shost_print_q(FILE *f, const shost_params *p, const shost_input *in,
                   const shost_results *r);                        /* output.c:831-912 */
void shost_print_p(FILE *f, const shost_params *p, const shost_input *in,
                   const shost_results *r);                        /* output.c:956-1005 */
double shost_est_log_prob(double sumlikes, double sumsqlikes, int reps);   /* output.c:457-467 */

/* INTERMEDSAVE's save interval (Init.c:377-385): NUMREPS / (INTERMEDSAVE + 1),
 * rounded up; 0 without INTERMEDSAVE. */
int32_t shost_save_freq(const shost_params *p);

/* CalcLike (CalcLike.c:387-426) of (Q [N][K], P [L][K][A]) on the recoded
 * genotypes, on the host: PRINTLIKES prints it for burn-in sweeps, whose lnP(D)
 * the sweep does not accumulate (PrintLike, output.c:187-200). */
double shost_calc_like(const shost_input *in, int K, const double *Q, const double *P);

#ifdef __cplusplus
}
#endif
#endif
