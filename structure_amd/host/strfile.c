/* Genotype data file of the `structure` command and its allele recoding.
 *
 * File layout (ReadData, datain.c:62-333): optional rows of marker names,
 * recessive alleles and map distances; then per individual PLOIDY rows (or
 * one row with PLOIDY columns per locus, ONEROWPERIND) of: label, population,
 * pop flag, location, phenotype / extra columns as switched on in mainparams,
 * then NUMLOCI integer alleles.  Tokens are separated by spaces, tabs or
 * newlines.  Alleles are then recoded per locus to 0..k-1 in order of first
 * appearance (individual-major, copy, CountAlleles datain.c:824-904), with a
 * translation table back to the file's values. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sgpu.h"
#include "../../include/structure_host.h"

typedef struct {
    FILE *f;
} tok_t;

static int is_white(int c) { return c == ' ' || c == '\n' || c == '\t'; }

/* Next token, at most max-1 characters (a longer token is cut, the rest
 * skipped).  Returns its length, 0 at end of file. */
static int token(tok_t *t, char *w, int max)
{
    int n = 0, c;
    while ((c = fgetc(t->f)) != EOF) {
        if (!is_white(c)) {
            if (n < max - 1) {
                w[n++] = (char)c;
            } else {
                while ((c = fgetc(t->f)) != EOF && !is_white(c)) {
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

static int valid_int(const char *s, int ind, int loc)
{
    for (const char *c = s; *c; c++) {
        if ((*c < '0' || *c > '9') && *c != '-') {
            printf("\nWARNING! Probable error in the input file.  \n");
            printf("Individual %d, locus %d;  ", ind + 1, loc + 1);
            printf("encountered the following data \n");
            printf("\"%s\" when expecting an integer\n\n", s);
            return 0;
        }
    }
    return 1;
}

static int eof_error(int ind)
{
    printf("\n\nWARNING:  Unexpected end of input file.  The details of the\n");
    printf("input file are set in mainparams.  I ran out of data while reading\n");
    printf("the data for individual %d.\n\n", ind + 1);
    return SGPU_ERR_ARG;
}

void shost_free_input(shost_input *in)
{
    free(in->geno);
    free(in->nall);
    free(in->translation);
    free(in->label);
    free(in->population);
    free(in->popflag);
    free(in->phenotype);
    free(in->location);
    free(in->markername);
    free(in->mapdistance);
    memset(in, 0, sizeof(*in));
}

/* Per-individual leading columns of one row. */
typedef struct {
    char label[SHOST_LABELLEN];
    int pop, popflag, location, phen;
} lead_t;

static int read_lead(tok_t *t, const shost_params *p, int ind, int loc, lead_t *o, int *valid)
{
    char s[SHOST_ALLELLEN];
    memset(o, 0, sizeof(*o));
    o->pop = o->popflag = o->location = o->phen = -1;
    if (p->label && !token(t, o->label, SHOST_LABELLEN)) {
        return eof_error(ind);
    }
    if (p->popdata) {
        if (!token(t, s, SHOST_ALLELLEN)) {
            return eof_error(ind);
        }
        *valid &= valid_int(s, ind, loc);
        o->pop = atoi(s);
    }
    if (p->popflag) {
        if (!token(t, s, SHOST_ALLELLEN)) {
            return eof_error(ind);
        }
        *valid &= valid_int(s, ind, loc);
        o->popflag = atoi(s);
    }
    if (p->locdata == 1) {
        if (!token(t, s, SHOST_ALLELLEN)) {
            return eof_error(ind);
        }
        valid_int(s, ind, loc);   /* the reference adds rather than multiplies here */
        o->location = atoi(s);
    }
    int phencol = -1;
    if (p->phenotype) {
        const int before = p->label + p->popdata + p->popflag;
        if (p->phenotypecol == -9) {
            phencol = 0;
        } else if (p->phenotypecol <= before || p->phenotypecol > before + 1 + p->extracols) {
            printf("Error: PHENOTYPECOL (%d) must lie in [%d, %d]\n", p->phenotypecol,
                   before + 1, before + 1 + p->extracols);
            return SGPU_ERR_ARG;
        } else {
            phencol = p->phenotypecol - before - 1;
        }
    }
    const int ncols = p->extracols + (p->phenotype ? 1 : 0);
    for (int col = 0; col < ncols; col++) {
        const int n = token(t, s, SHOST_ALLELLEN);
        if (col == phencol) {
            if (!n) {
                return eof_error(ind);
            }
            *valid &= valid_int(s, ind, loc);
            o->phen = atoi(s);
        }
    }
    return SGPU_OK;
}

/* First-appearance recoding (CountAlleles, datain.c:824-904). */
static int recode(shost_params *p, shost_input *in)
{
    const int N = in->N, L = in->L, lines = in->lines;
    /* a missing code that could collide with a recoded allele becomes -9 */
    const int newmissing = (p->missing >= 0 && p->missing < in->maxalleles + 1) ? -9 : p->missing;
    in->translation = (int32_t *)calloc((size_t)L * (in->maxalleles > 0 ? in->maxalleles : 1),
                                        sizeof(int32_t));
    in->nall = (int32_t *)calloc((size_t)L, sizeof(int32_t));
    if (!in->translation || !in->nall) {
        return SGPU_ERR_NOMEM;
    }
    for (int l = 0; l < L; l++) {
        int32_t *tr = in->translation + (size_t)l * in->maxalleles;
        int k = 0;
        for (int i = 0; i < N; i++) {
            for (int c = 0; c < lines; c++) {
                int32_t *g = in->geno + ((size_t)i * lines + c) * L + l;
                if (*g == p->missing) {
                    *g = newmissing;
                    continue;
                }
                int a = 0;
                while (a < k && tr[a] != *g) {
                    a++;
                }
                if (a == k) {
                    tr[k++] = *g;
                }
                *g = a;
            }
        }
        in->nall[l] = k;
        if (k == 0) {
            printf("WARNING: locus %d has all missing data.  This might cause unexpected "
                   "behaviour.\n",
                   l);
        }
    }
    in->missing = newmissing;
    return SGPU_OK;
}

/* Number of distinct alleles per locus: min, mean, max (FindMaxAlleles,
 * datain.c:768-820, before recoding). */
static int count_alleles(const shost_params *p, const shost_input *in)
{
    const int N = in->N, L = in->L, lines = in->lines;
    int32_t *seen = (int32_t *)malloc(sizeof(int32_t) * (size_t)(N * lines + 1));
    if (!seen) {
        return -1;
    }
    int mink = lines * N + 1, maxk = 0;
    long sumk = 0;
    for (int l = 0; l < L; l++) {
        int k = 0;
        for (int i = 0; i < N; i++) {
            for (int c = 0; c < lines; c++) {
                const int32_t v = in->geno[((size_t)i * lines + c) * L + l];
                if (v == p->missing) {
                    continue;
                }
                int a = 0;
                while (a < k && seen[a] != v) {
                    a++;
                }
                if (a == k) {
                    seen[k++] = v;
                }
            }
        }
        maxk = k > maxk ? k : maxk;
        mink = k < mink ? k : mink;
        sumk += k;
    }
    free(seen);
    printf("Number of alleles per locus: min=%2d; ave=%2.1f; max=%2d\n", mink,
           (double)sumk / L, maxk);
    return maxk;
}

static int read_input(shost_params *p, shost_input *in);

int shost_read_input(shost_params *p, shost_input *in)
{
    const int rc = read_input(p, in);
    fflush(stdout);   /* messages reach the terminal before the caller reacts */
    return rc;
}

static int read_input(shost_params *p, shost_input *in)
{
    memset(in, 0, sizeof(*in));
    if (p->ploidy != 2) {
        printf("PLOIDY=%d: only diploid data are supported\n", p->ploidy);
        return SGPU_ERR_UNSUPPORTED;
    }
    if (p->recessivealleles) {
        printf("RECESSIVEALLELES is not supported (the admixture sweep has no recessive model)\n");
        return SGPU_ERR_UNSUPPORTED;
    }
    FILE *f = fopen(p->datafile, "r");
    if (!f) {
        printf("Unable to open the file %s.\n", p->datafile);
        return SGPU_ERR_ARG;
    }
    printf("Reading file \"%s\".\n", p->datafile);
    tok_t t = {f};
    const int N = p->numinds, L = p->numloci, lines = p->ploidy;
    in->N = N;
    in->L = L;
    in->lines = lines;
    in->geno = (int32_t *)malloc(sizeof(int32_t) * (size_t)N * lines * L);
    in->label = (char *)calloc((size_t)N, SHOST_LABELLEN);
    in->population = (int32_t *)calloc((size_t)N, sizeof(int32_t));
    in->popflag = (int32_t *)calloc((size_t)N, sizeof(int32_t));
    in->phenotype = (int32_t *)calloc((size_t)N, sizeof(int32_t));
    in->location = (int32_t *)calloc((size_t)N, sizeof(int32_t));
    in->markername = (char *)calloc((size_t)L, SHOST_GENELEN);
    in->mapdistance = (double *)calloc((size_t)L, sizeof(double));
    if (!in->geno || !in->label || !in->population || !in->popflag || !in->phenotype ||
        !in->location || !in->markername || !in->mapdistance) {
        fclose(f);
        shost_free_input(in);
        return SGPU_ERR_NOMEM;
    }
    int rc = SGPU_OK, valid = 1;
    char s[SHOST_GENELEN > SHOST_ALLELLEN ? SHOST_GENELEN : SHOST_ALLELLEN];
    if (p->markernames) {
        for (int l = 0; l < L && rc == SGPU_OK; l++) {
            if (!token(&t, s, SHOST_GENELEN)) {
                printf("%i MARKERNAMES\n", l);
                rc = eof_error(-1);
            } else {
                memcpy(in->markername + (size_t)l * SHOST_GENELEN, s, strlen(s) + 1);
            }
        }
    }
    if (p->mapdistances) {
        for (int l = 0; l < L && rc == SGPU_OK; l++) {
            if (!token(&t, s, SHOST_ALLELLEN)) {
                rc = eof_error(-1);
            } else {
                in->mapdistance[l] = atof(s);
            }
        }
    }
    const int rows = p->onerowperind ? 1 : lines;
    const int cols = p->onerowperind ? lines : 1;
    for (int i = 0; i < N && rc == SGPU_OK; i++) {
        for (int row = 0; row < rows && rc == SGPU_OK; row++) {
            lead_t ld;
            if ((rc = read_lead(&t, p, i, L, &ld, &valid))) {
                break;
            }
            if (row == 0) {
                memcpy(in->label + (size_t)i * SHOST_LABELLEN, ld.label, SHOST_LABELLEN);
                in->population[i] = ld.pop;
                in->popflag[i] = ld.popflag;
                in->phenotype[i] = ld.phen;
                in->location[i] = p->locispop ? ld.pop : ld.location;
            }
            for (int l = 0; l < L && rc == SGPU_OK; l++) {
                for (int c = 0; c < cols; c++) {
                    if (!token(&t, s, SHOST_ALLELLEN)) {
                        printf("readlociEOF\n");
                        rc = eof_error(i);
                        break;
                    }
                    valid &= valid_int(s, i, l);
                    const int copy = p->onerowperind ? c : row;
                    in->geno[((size_t)i * lines + copy) * L + l] = atoi(s);
                }
            }
        }
        if (rc == SGPU_OK && p->phaseinfo) {   /* a row of phase probabilities: skipped */
            for (int l = 0; l < L; l++) {
                if (!token(&t, s, SHOST_ALLELLEN)) {
                    rc = eof_error(i);
                    break;
                }
            }
        }
    }
    if (rc == SGPU_OK) {
        int c;
        while ((c = fgetc(f)) != EOF) {
            if (!is_white(c)) {
                printf("\n\nWARNING:  There may be more data in the input file\n");
                printf("than indicated by the program constants.  Check the values\n");
                printf("entered for NUMLOCI and NUMINDS, etc, in the program constants.\n\n\n");
                valid = 0;
                break;
            }
        }
    }
    fclose(f);
    if (rc == SGPU_OK && !valid) {
        rc = SGPU_ERR_ARG;
    }
    if (rc == SGPU_OK) {
        in->maxalleles = count_alleles(p, in);
        if (in->maxalleles < 0) {
            rc = SGPU_ERR_NOMEM;
        }
    }
    if (rc == SGPU_OK) {
        p->origmissing = p->missing;
        rc = recode(p, in);
    }
    if (rc != SGPU_OK) {
        shost_free_input(in);
    }
    return rc;
}

int shost_geno_int8(const shost_input *in, int8_t *out)
{
    const size_t n = (size_t)in->N * in->lines * in->L;
    for (size_t j = 0; j < n; j++) {
        const int32_t g = in->geno[j];
        if (g == in->missing) {
            out[j] = -9;
        } else if (g < 0 || g > 126) {
            printf("allele code %d does not fit the packed genotype formats\n", g);
            return SGPU_ERR_UNSUPPORTED;
        } else {
            out[j] = (int8_t)g;
        }
    }
    return SGPU_OK;
}
