/*
 * sgpu.h -- C-ABI of the MI355X (gfx950) STRUCTURE admixture sweep.
 *
 * Drop-in boundary for the reference's device layer: the *CL host wrappers
 * and the OpenCL dispatch runtime of pmelsted/structure.  Each entry point
 * names the reference interface it replaces.  Plain C types only; host
 * buffers are caller-owned, device buffers library-owned.  Every function
 * returns SGPU_OK (0) or a negative status; nothing inside the library calls
 * exit() (the reference exits through handleCLErr, Kernels.c:195-204).
 * One HIP stream per context; a context is not thread-safe.
 */
#ifndef SGPU_H
#define SGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
enum {
    SGPU_OK = 0,
    SGPU_ERR_ARG = -1,          /* bad argument / shape */
    SGPU_ERR_NOMEM = -2,        /* device or host allocation failed */
    SGPU_ERR_DEVICE = -3,       /* HIP runtime error (no gfx950 device, launch failure) */
    SGPU_ERR_STATE = -4,        /* call out of order (e.g. sweep before init) */
    SGPU_ERR_UNSUPPORTED = -5   /* option or data the build does not handle */
};

/* ---- genotype formats --------------------------------------------------- */
enum {
    /* Biallelic, 2 bits per (individual, locus): 0 = (0,0), 1 = (0,1), 2 = (1,1),
     * 3 = both copies missing.  Words are uint32 laid out [ceil(N/16)][Lpad]:
     * word (g, l) holds individuals 16g..16g+15 at locus l, individual 16g+j in
     * bits 2j..2j+1.  Lpad = L rounded up to 256; padding codes are 3. */
    SGPU_GENO_SNP2 = 1,
    /* Any allele count, one byte per allele copy (255 = missing).  Words are
     * uint32 laid out [L][Gpad], locus-major: word (h, l) at l * Gpad + h holds
     * bytes (ind 2h copy 0, ind 2h copy 1, ind 2h+1 copy 0, ind 2h+1 copy 1), low
     * byte first.  Gpad = ceil(N/2) rounded up to 64; padding words are all 255.
     * Allele indices are < 255 (MAXALLELES <= 255; sgpu_pack_byte32). */
    SGPU_GENO_BYTE = 2
};

/* Model options: the mainparams/extraparams keys the admixture sweep reads
 * (structure.h:70-131; defaults params.c:571-670).  Replaces the -D constants
 * baked into the OpenCL JIT (Kernels.c:722-732) with runtime values. */
typedef struct {
    int32_t N, L, K, A;           /* NUMINDS, NUMLOCI, MAXPOPS, MAXALLELES */
    int32_t burnin, numreps;      /* BURNIN, NUMREPS */
    int32_t freqscorr, onefst;    /* FREQSCORR, ONEFST */
    int32_t inferalpha, popalphas, unifprioralpha;
    int32_t inferlambda, popspecificlambda;
    int32_t metrofreq, computeprob;
    double alpha0, lambda0;       /* ALPHA, LAMBDA */
    double alphapropsd, alphamax, alphapriora, alphapriorb;
    double fpriormean, fpriorsd;
    uint32_t seed, chain;         /* Philox key */
    /* ANCESTDIST, NUMBOXES: histogram of each Q[i,k] over the collected sweeps in
     * NUMBOXES equal boxes of [0,1] (UpdateSums, output.c:691-700); read back as
     * SGPU_ANCESTDIST.  NUMBOXES is ignored when ANCESTDIST is 0. */
    int32_t ancestdist, numboxes;
    /* PFROMPOPFLAGONLY: the allele counts behind the P update take only the
     * individuals whose POPFLAG is 1 (GetNumFromPops, UpdateP.c:85-88); the
     * ancestry counts behind Q still take every individual.  Needs
     * sgpu_data.popflag. */
    int32_t pfrompopflagonly;
    int32_t reserved;             /* 0 */
} sgpu_model;

typedef struct {
    int32_t format;               /* SGPU_GENO_SNP2 or SGPU_GENO_BYTE */
    const uint32_t *geno;         /* packed words (host memory) */
    size_t geno_words;            /* number of uint32 words */
    const int32_t *nall;          /* [L] NumAlleles after recoding (datain.c:824-904) */
    /* [N] POPFLAG column of the data file (Individual[i].PopFlag, datain.c), or
     * NULL when the file has none; read only by chains with PFROMPOPFLAGONLY. */
    const int32_t *popflag;
} sgpu_data;

typedef struct sgpu_ctx sgpu_ctx;

/* ---- state selectors for sgpu_get_state / sgpu_set_state ----------------- */
enum {
    SGPU_Q = 1,          /* double [N][K]            Q (structure.h:29)          */
    SGPU_P = 2,          /* double [L][K][A]         P (structure.h:28)          */
    SGPU_EPS = 3,        /* double [L][A]            Epsilon                     */
    SGPU_FST = 4,        /* double [K]               Fst                         */
    SGPU_ALPHA = 5,      /* double [K]               Alpha                       */
    SGPU_LAMBDA = 6,     /* double [K]               lambda                      */
    SGPU_NA = 7,         /* int32  [L][K][A]         NumAFromPops (UpdateP.c:115) */
    SGPU_NQ = 8,         /* int32  [N][K]            NumLociPops (structure.c:417) */
    SGPU_QSUM = 9,       /* double [N][K]            */
    SGPU_PSUM = 10,      /* double [L][K][A]         */
    SGPU_FSTSUM = 11,    /* double [K]               */
    SGPU_EPSSUM = 12,    /* double [L][A]            */
    SGPU_ALPHASUM = 13,  /* double [K]               */
    SGPU_LAMBDASUM = 14, /* double [K]               */
    SGPU_LIKE = 15,      /* double [3]: like, sumlikes, sumsqlikes (output.c:86-99) */
    SGPU_NSUM = 16,      /* int64  [1]: number of collected sweeps */
    SGPU_ANCESTDIST = 17 /* int32  [N][K][NUMBOXES]  AncestDist (structure.h:35); ANCESTDIST=1 */
};

/* Flags for sgpu_update_z. */
enum {
    SGPU_Z_LIKE = 1      /* also accumulate lnP(D) of (Q, P) (CalcLike.c:387-426) */
};

/* ---- genotype packing (host only; datain.c:768-904 produce the recoded input) ---- */

/* Words needed for format fmt at N individuals and L loci. */
size_t sgpu_packed_words(int32_t fmt, int32_t N, int32_t L);

/* Pack recoded genotypes geno[N][2][L] (reference GenPos layout, structure.h:22;
 * -9 = missing) into SGPU_GENO_SNP2.  Returns SGPU_ERR_UNSUPPORTED when a site
 * is not representable (allele > 1, one copy missing, or heterozygote stored
 * as (1,0)); such data goes through SGPU_GENO_BYTE, which keeps file order. */
int sgpu_pack_snp2(const int8_t *geno, int32_t N, int32_t L, uint32_t *out, size_t out_words);
int sgpu_pack_byte(const int8_t *geno, int32_t N, int32_t L, uint32_t *out, size_t out_words);

/* SGPU_GENO_BYTE from the reference's own Geno array (int32 [N][2][L], GenPos,
 * structure.h:22) with its missing code: alleles 0..254, so loci with up to
 * 255 alleles (the int8 form above stops at 127). */
int sgpu_pack_byte32(const int32_t *geno, int32_t N, int32_t L, int32_t missing, uint32_t *out,
                     size_t out_words);

/* ---- context ------------------------------------------------------------ */

/* Replaces InitCLDict (Kernels.c:641-751): selects the current HIP device,
 * allocates device state and uploads the packed genotypes. */
int sgpu_create(const sgpu_model *m, const sgpu_data *d, sgpu_ctx **out);

/* Tile shapes of the batched launches; 0 selects the library default.
 *   z_blocks      blocks per 2-bit UpdateZ launch over the whole batch (sets the
 *                 individuals per block; default 4096)
 *   metro_blocks  blocks per Metropolis-ratio launch (sets the loci per block;
 *                 default 2048)
 *   byte_loci     loci per byte-format UpdateZ block (default: the longest of 64,
 *                 32, 16 that gives z_blocks blocks over the batch, else 16 to 64
 *                 by K; at most 256)
 *   byte_chunk    loci whose K x A frequencies and counts are staged in LDS at
 *                 once by the byte-format UpdateZ (default and upper bound: 16
 *                 loci within 32 KiB)
 * The draws and counts do not depend on them; lnP(D) and the Metropolis
 * log-ratios are FP64 sums whose grouping does (differences in the last bits).
 * structure_amd/tune.py searches them per shape. */
typedef struct {
    int32_t z_blocks;
    int32_t metro_blocks;
    int32_t byte_loci;
    int32_t byte_chunk;
} sgpu_tuning;

/* A batch of chains in one context (batched form of InitCLDict; the reference
 * runs one chain per process).  Chain c uses models[c] and data set
 * data[data_of_chain[c]]; chains sharing a data set must agree on N, L, A, and
 * all chains on BURNIN, METROFREQ and COMPUTEPROB.  K, the flags, the priors
 * and the Philox key (seed, chain) may differ per chain.  Every per-phase call
 * and sweep below then advances all chains with one launch per kernel (per
 * distinct (format, K) for UpdateZ and the Metropolis ratio).  Each chain's
 * random stream depends only on its own key, so a chain's Z, counts, P, Q,
 * alpha, lambda, eps and F are those of the same chain run alone.
 * tuning may be NULL. */
int sgpu_create_batch(const sgpu_model *models, int32_t nchains, const sgpu_data *data,
                      int32_t ndata, const int32_t *data_of_chain, const sgpu_tuning *tuning,
                      sgpu_ctx **out);

/* Number of chains in the context. */
int sgpu_num_chains(const sgpu_ctx *ctx);

/* Replaces ReleaseCLDict (Kernels.c:177-193). */
int sgpu_destroy(sgpu_ctx *ctx);

/* Initialization (Init.c:329-444) for the admixture model: alpha, lambda,
 * InitFreqPriors (Init.c:172-240), Q = 1/K (structure.c:51-58), random initial
 * Z (Init.c:124-169) reduced to its counts, posterior sums zeroed. */
int sgpu_init_state(sgpu_ctx *ctx);

/* Replace writeBuffer/readBuffer (Kernels.c:805,856).  bytes must equal the
 * size of the selected array.  get synchronizes the context stream. */
int sgpu_get_state(sgpu_ctx *ctx, int32_t which, void *host, size_t bytes);
int sgpu_set_state(sgpu_ctx *ctx, int32_t which, const void *host, size_t bytes);

/* The same for chain `chain` of a batch (sgpu_get_state reads chain 0). */
int sgpu_get_chain_state(sgpu_ctx *ctx, int32_t chain, int32_t which, void *host, size_t bytes);
int sgpu_set_chain_state(sgpu_ctx *ctx, int32_t chain, int32_t which, const void *host,
                         size_t bytes);

/* ---- per-phase updates (asynchronous on the context stream) --------------- */
/* Each call advances every chain of the context.  The alpha, lambda, epsilon
 * and Fst updates act on the chains whose model enables them (INFERALPHA,
 * INFERLAMBDA, FREQSCORR), as structure.c:581-605 guards their calls. */

/* UpdatePCL (UpdateP.c:202-273, UpdateP.h:4-7): P[l,k,:] ~ Dirichlet from the
 * allele counts of the last UpdateZ.  collect != 0 also adds P into PSum. */
int sgpu_update_p(sgpu_ctx *ctx, uint32_t rep, int32_t collect);

/* UpdateQCL (UpdateQ.c:894-929, UpdateQ.h:6-9): Metropolis from the prior when
 * metrofreq > 0 and rep % metrofreq == 0, else Gibbs from the ancestry counts. */
int sgpu_update_q(sgpu_ctx *ctx, uint32_t rep, int32_t collect);

/* UpdateZCL (UpdateZ.c:60-93, UpdateZ.h:2-3) fused with GetNumFromPops
 * (UpdateP.cl:1-41), GetNumLociPops (UpdateQ.cl:28-61) and, with SGPU_Z_LIKE,
 * the per-copy lnP(D) terms of mapReduceLogLike (calcLike.cl:204-259).  Z is not
 * materialised; zout (host, int8 [N][2][L], may be NULL) requests a debug copy. */
int sgpu_update_z(sgpu_ctx *ctx, uint32_t rep, int32_t flags, int8_t *zout);

/* UpdateAlphaCL (UpdateAlpha.c:307-393, UpdateAlpha.h:3-4). */
int sgpu_update_alpha(sgpu_ctx *ctx, uint32_t rep);

/* UpdateLambda / UpdatePopLambda (UpdateLambda.c:8-102, UpdateLambda.h:1-3),
 * host-only in the reference; here on the device, without the P readback. */
int sgpu_update_lambda(sgpu_ctx *ctx, uint32_t rep);

/* UpdateEpsilonCL (UpdateEpsilon.c:270-292, UpdateEpsilon.h:4-5), restated from
 * the CPU NonIndependenceUpdateEpsilon (UpdateEpsilon.c:71-145).  collect != 0
 * also adds Epsilon into SumEpsilon. */
int sgpu_update_eps(sgpu_ctx *ctx, uint32_t rep, int32_t collect);

/* UpdateFstCL (UpdateFst.c:100-140, UpdateFst.h:1). */
int sgpu_update_fst(sgpu_ctx *ctx, uint32_t rep);

/* DataCollectionCL (output.c:104-183): alpha/lambda/Fst sums and, when the last
 * UpdateZ ran with SGPU_Z_LIKE, like += ..., sumlikes, sumsqlikes. */
int sgpu_collect(sgpu_ctx *ctx, uint32_t rep);

/* One sweep of the hot loop (structure.c:497-638), with the reference's
 * draws and results.  The phases run in four launches (six with a Metropolis
 * Q update): the Gibbs Q update, alpha, lambda and the Fst terms ride in the
 * launches of phases whose inputs and outputs they do not touch. */
int sgpu_sweep(sgpu_ctx *ctx, uint32_t rep);

/* n sweeps rep0 .. rep0+n-1, the same draws and results as n calls of
 * sgpu_sweep, replayed from HIP graphs captured on first use: graphs of
 * several sweeps (whole Metropolis periods of up to 64 sweeps, else 16 Gibbs
 * sweeps) where a span lies wholly in burn-in or in the collecting sweeps,
 * single-sweep graphs around them. */
int sgpu_run_sweeps(sgpu_ctx *ctx, uint32_t rep0, uint32_t n);

/* ---- phase timing ------------------------------------------------------- */
/* The reference's queue is created without profiling (Kernels.c:691); these
 * record HIP events around every phase launch on the context stream.  In a
 * sweep, the Gibbs Q update counts under P (Metropolis Q updates under Q);
 * lambda's decision under P (F model) or EPS, the log sums of P it reads under
 * Q in Metropolis sweeps and LAMBDA in Gibbs sweeps (no F model), the log-eps
 * sums (F model), alpha and the Fst terms under EPS, the Fst decision under
 * COLLECT; ALPHA, LAMBDA and FST time the stand-alone calls. */
enum {
    SGPU_PH_P = 0, SGPU_PH_Q = 1, SGPU_PH_Z = 2, SGPU_PH_ALPHA = 3, SGPU_PH_LAMBDA = 4,
    SGPU_PH_EPS = 5, SGPU_PH_FST = 6, SGPU_PH_COLLECT = 7, SGPU_NPHASE = 8
};
int sgpu_set_timing(sgpu_ctx *ctx, int32_t enable);
/* Total milliseconds and launch counts per phase since the last call (then reset). */
int sgpu_get_timing(sgpu_ctx *ctx, double *ms, int64_t *launches);

/* Wait for the context stream (finishCommands, Kernels.c:781-790). */
int sgpu_synchronize(sgpu_ctx *ctx);

/* Diagnostics, no counterpart in the reference: y[i] = f(x[i]) evaluated by the
 * device code the sweep uses (host arrays in and out; a temporary device
 * allocation on the current device).  SGPU_FN_LOG: natural logarithm;
 * SGPU_FN_COS / SGPU_FN_SIN: cosine / sine for 0 <= x <= 8 (the Box-Muller
 * angle); SGPU_FN_LGAMMA: ln Gamma(x), x > 0, by the Lanczos sum of mylgamma
 * (mymath.c:382-429). */
enum { SGPU_FN_LOG = 0, SGPU_FN_COS = 1, SGPU_FN_LGAMMA = 2, SGPU_FN_SIN = 3 };
int sgpu_eval_math(int32_t fn, const double *x, double *y, int64_t n);

/* sgpu_eval_math without any device allocation: no counterpart in the
 * reference, for callers that cannot allocate device memory (harnesses that
 * forbid hipMalloc, allocation-free serving loops).  x and y are host arrays
 * of n <= SGPU_EVAL_MATH_NOALLOC_MAX values; larger n returns SGPU_ERR_ARG.
 * The values are staged in a fixed device array of the library (1 MiB per GPU,
 * held until the process ends) that every call on that GPU shares: calls that
 * overlapped on one GPU would race, so the library runs them one at a time (a
 * process-wide lock around each synchronous call).  Nothing in the staging
 * needs clearing: every value a call reads was written by that call.  Runs on
 * the current HIP device. */
#define SGPU_EVAL_MATH_NOALLOC_MAX 65536
int sgpu_eval_math_noalloc(int32_t fn, const double *x, double *y, int64_t n);

/* Message for the last failing call on this thread. */
const char *sgpu_last_error(void);

/* Library build information (device architecture, kernel variants). */
const char *sgpu_version(void);

#ifdef __cplusplus
}
#endif
#endif
