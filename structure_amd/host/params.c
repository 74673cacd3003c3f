/* mainparams / extraparams / command line of the `structure` command.
 *
 * Keeps the reference's file syntax (params.c:225-570): the reader looks for
 * '#', then the word "define", then a key and its value; everything else in
 * the file is commentary.  Keys are table-driven here; the value syntax per
 * key follows the reference (decimal ints, %i ints for the LOCPRIOR group and
 * SEED, doubles, and the two file names read to end of line or "//"). */
#include <ctype.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sgpu.h"
#include "../../include/structure_host.h"

#define MAXNAME 30   /* longest key (params.c:34) */

enum { K_INT, K_IINT, K_DBL, K_PATH };

typedef struct {
    const char *name;
    int kind;
    size_t off;
} key_t_;

#define I(n, f) {n, K_INT, offsetof(shost_params, f)}
#define II(n, f) {n, K_IINT, offsetof(shost_params, f)}
#define D(n, f) {n, K_DBL, offsetof(shost_params, f)}

/* Keys accepted by SetValue (params.c:349-568). */
static const key_t_ KEYS[] = {
    {"INFILE", K_PATH, offsetof(shost_params, datafile)},
    {"OUTFILE", K_PATH, offsetof(shost_params, outfile)},
    I("NUMINDS", numinds), I("NUMLOCI", numloci), I("MISSING", missing), I("LABEL", label),
    I("POPDATA", popdata), I("MARKERNAMES", markernames), I("PHASEINFO", phaseinfo),
    I("MAPDISTANCES", mapdistances), I("MARKOVPHASE", markovphase), I("PLOIDY", ploidy),
    I("ONEROWPERIND", onerowperind), I("RECESSIVEALLELES", recessivealleles),
    I("NOTAMBIGUOUS", notambiguous), I("POPFLAG", popflag), I("PHENOTYPE", phenotype),
    I("EXTRACOLS", extracols), I("MAXPOPS", maxpops), I("BURNIN", burnin),
    I("NUMREPS", numreps), I("USEPOPINFO", usepopinfo), I("PFROMPOPFLAGONLY", pfrompopflagonly),
    I("INFERALPHA", inferalpha), I("INFERLAMBDA", inferlambda),
    I("POPSPECIFICLAMBDA", popspecificlambda), I("POPALPHAS", popalphas),
    I("COMPUTEPROB", computeprob), I("NOADMIX", noadmix), I("ADMBURNIN", admburnin),
    I("PHASED", phased), I("PRINTQHAT", printqhat), I("PRINTQSUM", printqsum),
    I("UPDATEFREQ", updatefreq), I("PRINTLIKES", printlikes), I("INTERMEDSAVE", intermedsave),
    I("PRINTKLD", printkld), I("PRINTNET", printnet), I("PRINTLAMBDA", printlambda),
    I("ECHODATA", echodata), I("ANCESTDIST", ancestdist), I("NUMBOXES", numboxes),
    D("ANCESTPINT", ancestpint), I("GENSBACK", gensback), D("MIGRPRIOR", migrprior),
    D("ALPHA", alpha), D("LOG10RPROPSD", log10rpropsd), D("LOG10RMIN", log10rmin),
    D("LOG10RMAX", log10rmax), D("LOG10RSTART", log10rstart), I("FREQSCORR", freqscorr),
    I("ONEFST", onefst), D("FPRIORMEAN", fpriormean), D("FPRIORSD", fpriorsd),
    D("LAMBDA", lambda), I("UNIFPRIORALPHA", unifprioralpha), D("ALPHAMAX", alphamax),
    D("ALPHAPRIORA", alphapriora), D("ALPHAPRIORB", alphapriorb),
    D("ALPHAPROPSD", alphapropsd), I("STARTATPOPINFO", startatpopinfo),
    I("RANDOMIZE", randomize), I("LINKAGE", linkage), I("METROFREQ", metrofreq),
    I("REPORTHITRATE", reporthitrate), I("NOQS", noqs), I("POSTERIOR", posterior),
    I("PICTUREFILE", picturefile), I("INDIVIDUALR", individualr),
    I("NUMSIMSTATS", numsimstats), D("EMERROR", emerror), I("POOLFREQ", poolfreq),
    I("LOCUSxONLY", locusxonly), I("PHENOTYPECOL", phenotypecol),
    I("MISSINGPHENO", missingpheno), II("LOCPRIOR", locprior),
    II("UPDATELOCPRIOR", updatelocprior), D("LOCPRIORINIT", locpriorinit),
    D("MAXLOCPRIOR", maxlocprior), II("LOCISPOP", locispop), II("LOCDATA", locdata),
    D("LOCPRIORSTEP", locpriorstep), II("SEED", seed),
};

void shost_preset(shost_params *p)
{
    memset(p, 0, sizeof(*p));
    strcpy(p->outfile, "output");
    p->markovphase = -1;
    p->log10rmin = -8;
    p->log10rmax = 2;
    p->log10rstart = -4;
    p->log10rpropsd = 0.1;
    p->missing = -9;
    p->label = p->popdata = p->popflag = p->phenotype = p->extracols = -1;
    p->ploidy = 2;
    p->notambiguous = -939;
    p->maxpops = -1;
    p->burnin = 10000;
    p->numreps = 20000;
    p->inferalpha = 1;
    p->computeprob = 1;
    p->admburnin = p->burnin / 4;
    p->phased = 1;
    p->updatefreq = 200;
    p->numboxes = 1000;
    p->ancestpint = 0.90;
    p->gensback = 2;
    p->migrprior = 0.05;
    p->alpha = 1.0;
    p->freqscorr = 1;
    p->fpriormean = 0.1;
    p->fpriorsd = 0.1;
    p->lambda = 1.0;
    p->unifprioralpha = 1;
    p->alphamax = 10.0;
    p->alphapriora = 1.0;
    p->alphapriorb = 2.0;
    p->alphapropsd = 0.05;
    p->randomize = 1;
    p->metrofreq = 20;
    p->numsimstats = 1000;
    p->emerror = 0.001;
    p->poolfreq = 10;
    p->phenotypecol = -9;
    p->missingpheno = -9;
    p->posterior = 1;
    p->updatelocprior = 1;
    p->locpriorinit = 1.0;
    p->locpriorstep = 0.1;
    p->maxlocprior = 20.0;
    p->printlocprior = 1;
    p->seed = -1;
    p->origmissing = p->missing;
}

/* Next whitespace-delimited word (space, tab, newline), at most max-1 chars;
 * a longer word is cut and the rest skipped.  Returns its length (0 at EOF). */
static int next_word(FILE *f, char *w, int max)
{
    int n = 0, c;
    while ((c = fgetc(f)) != EOF) {
        const int white = c == ' ' || c == '\n' || c == '\t';
        if (!white) {
            if (n < max - 1) {
                w[n++] = (char)c;
            } else {
                /* too long: skip to the next whitespace */
                while ((c = fgetc(f)) != EOF && !(c == ' ' || c == '\n' || c == '\t')) {
                }
                break;
            }
        } else if (n > 0) {
            break;
        }
    }
    w[n] = '\0';
    return n;
}

/* A file name value: leading blanks skipped, read to end of line or "//",
 * trailing blanks dropped (spaces inside the name are kept). */
static void path_value(FILE *f, char *out, int max)
{
    int n = 0, c, last = 0;
    while (n < max - 1 && (c = fgetc(f)) != EOF) {
        if (last == '/' && c == '/') {
            n--;
            break;
        }
        if (n == 0 && isspace(c)) {
            continue;
        }
        if (c == '\n' || c == '\r') {
            break;
        }
        out[n++] = (char)c;
        last = c;
    }
    while (n > 0 && isspace((unsigned char)out[n - 1])) {
        n--;
    }
    out[n] = '\0';
}

static void set_key(shost_params *p, const char *key, FILE *f)
{
    for (size_t k = 0; k < sizeof(KEYS) / sizeof(KEYS[0]); k++) {
        if (strcmp(KEYS[k].name, key) != 0) {
            continue;
        }
        char *base = (char *)p + KEYS[k].off;
        switch (KEYS[k].kind) {
        case K_PATH:
            path_value(f, base, SHOST_STRLEN);
            if (base == p->datafile) {
                printf("datafile is\n%s\n", p->datafile);
            }
            break;
        case K_INT:
            if (fscanf(f, "%d", (int32_t *)base) != 1) {
                /* value left unchanged, as fscanf leaves it */
            }
            break;
        case K_IINT:
            if (fscanf(f, "%i", (int32_t *)base) != 1) {
            }
            break;
        default:
            if (fscanf(f, "%lf", (double *)base) != 1) {
            }
            break;
        }
        return;
    }
    printf("Warning:  Ignoring unrecognized parameter name (%s)\n", key);
}

int shost_read_params(shost_params *p, const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) {
        printf("Can't open the file \"%s\".\n", path);
        return SGPU_ERR_ARG;
    }
    printf("Reading file \"%s\".\n", path);
    int c;
    char w[MAXNAME];
    while ((c = fgetc(f)) != EOF) {
        if (c != '#') {
            continue;
        }
        if (next_word(f, w, MAXNAME) && strcmp(w, "define") == 0 && next_word(f, w, MAXNAME)) {
            set_key(p, w, f);
        }
    }
    fclose(f);
    p->origmissing = p->missing;
    return SGPU_OK;
}

void shost_param_file_name(int argc, char **argv, const char *flag, char *out, size_t n)
{
    out[0] = '\0';
    int found = 0;
    for (int i = 0; i < argc - 1 && !found; i++) {
        if (strcmp(argv[i], flag) != 0) {
            continue;
        }
        found = 1;
        /* the words up to the next flag, joined by spaces (names with spaces) */
        size_t len = 0;
        for (int j = i + 1; j < argc && argv[j][0] != '-'; j++) {
            const size_t wl = strlen(argv[j]);
            if (len + wl + 2 > n) {
                break;
            }
            if (len) {
                out[len++] = ' ';
            }
            memcpy(out + len, argv[j], wl);
            len += wl;
            out[len] = '\0';
        }
    }
    if (!found) {
        const char *def = strcmp(flag, "-m") == 0   ? "mainparams"
                          : strcmp(flag, "-e") == 0 ? "extraparams"
                                                    : "STRATparams";
        snprintf(out, n, "%s", def);
    }
    /* quotes delimit: drop an opening quote, stop at the closing one */
    char *q = strchr(out, '"');
    if (q) {
        memmove(q, q + 1, strlen(q));
        char *e = strchr(q, '"');
        if (e) {
            *e = '\0';
        }
    }
}

static int cl_int(const char *v, const char *name, int32_t *out)
{
    for (const char *s = v; *s; s++) {
        if ((*s < '0' || *s > '9') && *s != '-') {
            printf("Error in command-line data entry for %s\n", name);
            printf("Value input at command line is %s\n", v);
            return SGPU_ERR_ARG;
        }
    }
    *out = atoi(v);
    return SGPU_OK;
}

int shost_command_line(shost_params *p, int argc, char **argv)
{
    int rc = SGPU_OK;
    for (int i = 0; i < argc - 1 && rc == SGPU_OK; i++) {
        const char *a = argv[i], *v = argv[i + 1];
        if (strcmp(a, "-i") == 0) {
            snprintf(p->datafile, sizeof(p->datafile), "%s", v);
        } else if (strcmp(a, "-o") == 0) {
            snprintf(p->outfile, sizeof(p->outfile), "%s", v);
        } else if (strcmp(a, "-K") == 0) {
            rc = cl_int(v, "MAXPOPS", &p->maxpops);
        } else if (strcmp(a, "-L") == 0) {
            rc = cl_int(v, "NUMLOCI", &p->numloci);
        } else if (strcmp(a, "-N") == 0) {
            rc = cl_int(v, "NUMINDS", &p->numinds);
        } else if (strcmp(a, "-r") == 0) {
            rc = cl_int(v, "NUMSIMSTATS", &p->numsimstats);
        } else if (strcmp(a, "-c") == 0) {
            rc = cl_int(v, "PHENOTYPECOL", &p->phenotypecol);
        } else if (strcmp(a, "-D") == 0) {
            p->seed = atoi(v);
            i++;
        }
    }
    return rc;
}

int shost_check(const shost_params *p)
{
    int bad = 0;
    if (p->datafile[0] == '\0') {
        printf("Need to specify the name of the data file\n");
        bad = 1;
    }
    if (p->numinds < 1) {
        printf("Need to set NUMINDS>0 in the file mainparams\n");
        bad = 1;
    }
    if (p->numloci < 1) {
        printf("Need to set NUMLOCI>0 in the file mainparams\n");
        bad = 1;
    }
    if (p->label == -1) {
        printf("Need to set LABEL=0 or 1 in the file mainparams\n");
        bad = 1;
    }
    if (p->popdata == -1) {
        printf("Need to set POPDATA=0 or 1 in the file mainparams\n");
        bad = 1;
    }
    if (p->popflag == -1) {
        printf("Need to set POPFLAG=0 or 1 in the file mainparams\n");
        bad = 1;
    }
    if (p->phenotype == -1) {
        printf("Need to set PHENOTYPE=0 or 1 in the file mainparams\n");
        bad = 1;
    }
    if (p->extracols < 0) {
        printf("Need to set EXTRACOLS>=0 in the file mainparams\n");
        bad = 1;
    }
    if (p->locprior && p->locdata < 0 && (p->locispop == 0 || p->popdata < 0)) {
        printf("Need to specify location data when LOCPRIOR=%i\n", p->locprior);
        printf("LOCISPOP=%i POPDATA=%i\n", p->locispop, p->popdata);
        bad = 1;
    }
    if (p->locprior && p->linkage) {
        printf("Cannot run LOCPRIOR and the LINKAGE model together\n");
        bad = 1;
    }
    return bad ? SGPU_ERR_ARG : SGPU_OK;
}

int32_t shost_update_freq(const shost_params *p)
{
    if (p->updatefreq > 0) {
        return p->updatefreq;
    }
    /* about one line per 4e6 genotypes, rounded to a readable step */
    const int numgens = p->ploidy * p->numinds * p->numloci;
    const int f = numgens > 0 ? 4000000 / numgens : 0;
    if (f <= 15) {
        return 10;
    }
    if (f <= 37) {
        return 25;
    }
    if (f <= 75) {
        return 50;
    }
    if (f <= 250) {
        return 100 * ((f + 50) / 100);
    }
    return 500 * ((f + 250) / 500);
}

static int get_params(shost_params *p, int argc, char **argv);

int shost_get_params(shost_params *p, int argc, char **argv)
{
    const int rc = get_params(p, argc, argv);
    fflush(stdout);   /* messages reach the terminal before the caller reacts */
    return rc;
}

static int get_params(shost_params *p, int argc, char **argv)
{
    char name[500];
    int rc;
    shost_preset(p);
    shost_param_file_name(argc, argv, "-m", name, sizeof(name));
    if ((rc = shost_read_params(p, name))) {
        return rc;
    }
    shost_param_file_name(argc, argv, "-e", name, sizeof(name));
    if ((rc = shost_read_params(p, name))) {
        return rc;
    }
    if ((rc = shost_command_line(p, argc, argv)) || (rc = shost_check(p))) {
        return rc;
    }
    p->origmissing = p->missing;
    p->updatefreq = shost_update_freq(p);
    return SGPU_OK;
}

void shost_print_all_params(FILE *f, const shost_params *p)
{
    fprintf(f, "Values of parameters used in structure:\n");
    fprintf(f, "DATAFILE=%s,\t", p->datafile);
    fprintf(f, "OUTFILE=%s,\t", p->outfile);
    const struct {
        const char *n;
        int32_t v;
    } a[] = {{"NUMINDS", p->numinds}, {"NUMLOCI", p->numloci}, {"MISSING", p->origmissing},
             {"LABEL", p->label}, {"POPDATA", p->popdata}, {"POPFLAG", p->popflag},
             {"PHENOTYPE", p->phenotype}, {"EXTRACOLS", p->extracols}, {"MAXPOPS", p->maxpops},
             {"BURNIN", p->burnin}, {"NUMREPS", p->numreps}, {"USEPOPINFO", p->usepopinfo},
             {"INFERALPHA", p->inferalpha}, {"INFERLAMBDA", p->inferlambda},
             {"POPSPECIFICLAMBDA", p->popspecificlambda}, {"POPALPHAS", p->popalphas},
             {"COMPUTEPROB", p->computeprob}, {"NOADMIX", p->noadmix},
             {"ADMBURNIN", p->admburnin}, {"UPDATEFREQ", p->updatefreq},
             {"PRINTLIKES", p->printlikes}, {"INTERMEDSAVE", p->intermedsave},
             {"PRINTKLD", p->printkld}, {"PRINTNET", p->printnet},
             {"PRINTLAMBDA", p->printlambda}, {"ANCESTDIST", p->ancestdist},
             {"NUMBOXES", p->numboxes}};
    for (size_t i = 0; i < sizeof(a) / sizeof(a[0]); i++) {
        fprintf(f, "%s=%d,\t", a[i].n, a[i].v);
    }
    fprintf(f, "ANCESTPINT=%1.5f,\t", p->ancestpint);
    fprintf(f, "GENSBACK=%d,\t", p->gensback);
    fprintf(f, "MIGRPRIOR=%1.5f,\t", p->migrprior);
    fprintf(f, "PRINTQHAT=%d,\t", p->printqhat);
    fprintf(f, "PRINTQSUM=%d,\t", p->printqsum);
    fprintf(f, "ALPHA=%1.4f,\t", p->alpha);
    fprintf(f, "FREQSCORR=%d,\t", p->freqscorr);
    fprintf(f, "FPRIORMEAN=%1.4f,\t", p->fpriormean);
    fprintf(f, "FPRIORSD=%1.4f,\t", p->fpriorsd);
    fprintf(f, "ONEFST=%d,\t", p->onefst);
    fprintf(f, "LAMBDA=%1.4f,\t", p->lambda);
    fprintf(f, "UNIFPRIORALPHA=%d,\t", p->unifprioralpha);
    fprintf(f, "ALPHAMAX=%1.4f,\t", p->alphamax);
    fprintf(f, "ALPHAPRIORA=%1.4f,\t", p->alphapriora);
    fprintf(f, "ALPHAPRIORB=%1.4f,\t", p->alphapriorb);
    fprintf(f, "ALPHAPROPSD=%1.4f,\t", p->alphapropsd);
    fprintf(f, "STARTATPOPINFO=%d,\t", p->startatpopinfo);
    fprintf(f, "RANDOMIZE=%d,\t", p->randomize);
    fprintf(f, "LINKAGE=%d,\t", p->linkage);
    fprintf(f, "METROFREQ=%d,\t", p->metrofreq);
    fprintf(f, "REPORTHITRATE=%d,\t", p->reporthitrate);
    fprintf(f, "MARKOVPHASE=%d,\t", p->markovphase);
    fprintf(f, "PHASED=%d,\t", p->phased);
    fprintf(f, "PLOIDY=%d,\t", p->ploidy);
    fprintf(f, "PHASEINFO=%d\t", p->phaseinfo);
    fprintf(f, "LOCPRIOR=%d, \t", p->locprior);
    fprintf(f, "LOCPRIORINIT=%f, \t", p->locpriorinit);
    fprintf(f, "LOCDATA=%d, \t", p->locdata);
    fprintf(f, "LOCISPOP=%d, \t", p->locispop);
    fprintf(f, "LOCPRIORSTEP=%f, \t", p->locpriorstep);
    fprintf(f, "MAXLOCPRIOR=%f, \t", p->maxlocprior);
    fprintf(f, "SEED=%d,\t", p->seed);
    fprintf(f, "\n[STRAT parameters]:    ");
    fprintf(f, "NUMSIMSTATS=%d,\t", p->numsimstats);
    fprintf(f, "PHENOTYPECOL=%d,\t", p->phenotypecol);
    fprintf(f, "POOLFREQ=%d,\t", p->poolfreq);
    fprintf(f, "LOCUSxONLY=%d,\t", p->locusxonly);
    fprintf(f, "EMERROR=%1.5f,\t", p->emerror);
    fprintf(f, "MISSINGPHENO=%d,\t", p->missingpheno);
}
