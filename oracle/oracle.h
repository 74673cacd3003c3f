/*
 * CPU oracle for the STRUCTURE admixture sweep (test infrastructure only).
 *
 * This is a plain-C, FP64 restatement of the reference CPU update functions
 * (the Updates directory, CalcLike.c, output.c, Init.c, mymath.c, ran.c of
 * pmelsted/structure), with every random draw taken from the counter-based
 * Philox4x32-10 stream that the HIP kernels also use.  It exists only to
 * check the product: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may call it, nothing in structure_amd/ may.
 *
 * Random stream (shared contract with the device code, restated on both
 * sides independently):
 *   key     = (seed, chain)
 *   counter = (c0, c1, c2, c3) with c3 = (phase << 24) | attempt
 *   phase  1 UpdateZ     c0=rep  c1=locus  c2=ind>>1   word 2*(ind&1)+j for the
 *                        j-th non-missing copy of (ind, locus)
 *   phase  2 InitZ       c0=0    c1=locus  c2=ind>>1   word as for phase 1
 *   phase  3 UpdateP     c0=rep  c1=locus  c2=(pop<<16)|allele   attempt j
 *   phase  4 UpdateQ     c0=rep  c1=ind    c2=pop                attempt j
 *   phase  5 Q-Metro     c0=rep  c1=ind    c2=pop (proposal) / 0xffffffff (accept, word 0)
 *   phase  6 alpha       c0=rep  c1=pop    c2=0   words 0,1 normal, 2 accept
 *   phase  7 lambda      c0=rep  c1=pop    c2=0   words 0,1 normal, 2 accept
 *   phase  8 epsilon     c0=rep  c1=locus  c2=0   words 0 a1, 1 a2, 2 step, 3 accept
 *   phase  9 Fst         c0=rep  c1=pop    c2=0   words 0,1 normal, 2 accept
 * Uniforms: uz(w) = w * 2^-32 in [0,1) (categorical draws, RandomInteger,
 * RandomReal); uo(w) = (w + 0.5) * 2^-32 in (0,1) (rnd(): acceptance tests,
 * Box-Muller, Marsaglia-Tsang).  Normal = Box-Muller cos branch on
 * (uo(w0), uo(w1)).  Gamma = Marsaglia-Tsang (ACM TOMS 26(3) 2000), one
 * Philox call per attempt, shape < 1 boosted by uo(w3)^(1/a) of the
 * accepted attempt.
 */
#ifndef STRUCTURE_ORACLE_H
#define STRUCTURE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_PH_Z = 1, ORC_PH_INITZ = 2, ORC_PH_P = 3, ORC_PH_Q = 4, ORC_PH_QMETRO = 5,
    ORC_PH_ALPHA = 6, ORC_PH_LAMBDA = 7, ORC_PH_EPS = 8, ORC_PH_FST = 9
};

#define ORC_MISSING (-9)    /* recoded missing allele, structure.h:8 UNASSIGNED / datain.c:838 */
#define ORC_UNASSIGNED (-9)

/* Model options: the subset of mainparams/extraparams the admixture sweep reads. */
typedef struct {
    int N, L, K, A;            /* individuals, loci, clusters (MAXPOPS), MAXALLELES stride */
    int burnin, numreps;
    int freqscorr, onefst;
    int inferalpha, popalphas, unifprioralpha;
    int inferlambda, popspecificlambda;
    int metrofreq, computeprob;
    double alpha0, lambda0;
    double alphapropsd, alphamax, alphapriora, alphapriorb;
    double fpriormean, fpriorsd;
    uint32_t seed, chain;
    int ancestdist, numboxes;   /* ANCESTDIST histogram of Q (output.c:691-700) */
    int pfrompopflagonly;       /* allele counts from POPFLAG=1 only (UpdateP.c:85-88) */
    const int32_t *popflag;     /* [N] Individual[i].PopFlag, or NULL */
} orc_model;

/* Chain state.  Layouts follow structure.h:21-33 (GenPos, PPos, QPos, EpsPos);
 * genotypes are int16 [N][2][L] (alleles 0..254, ORC_MISSING for a missing copy). */
typedef struct {
    double *Q;        /* [N][K]   */
    double *P;        /* [L][K][A] */
    double *eps;      /* [L][A]   */
    double *F;        /* [K]      */
    double *alpha;    /* [K]      */
    double *lambda;   /* [K]      */
    int32_t *nA;      /* [L][K][A] NumAFromPops */
    int32_t *nQ;      /* [N][K]    NumLociPops  */
    double *QSum, *PSum, *FstSum, *epsSum, *alphaSum, *lambdaSum; /* posterior sums */
    double like, sumlikes, sumsqlikes;
    int64_t nsum;     /* number of collected sweeps */
    int32_t *ancestDist;   /* [N][K][NUMBOXES] (AncestDistPos, structure.h:35) when ANCESTDIST */
} orc_state;

/* --- RNG --------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_philox_block(uint32_t seed, uint32_t chain, uint32_t c0, uint32_t c1, uint32_t c2,
                      uint32_t c3, uint32_t out[4]);
double orc_uz(uint32_t w);
double orc_uo(uint32_t w);
double orc_normal(uint32_t seed, uint32_t chain, uint32_t c0, uint32_t c1, uint32_t c2,
                  uint32_t phase);
double orc_gamma(double a, uint32_t seed, uint32_t chain, uint32_t c0, uint32_t c1,
                 uint32_t c2, uint32_t phase);
double orc_lgamma(double z);

/* --- initialisation (Init.c:124-240, 350-414) --------------------------- */
void orc_init_state(const orc_model *m, const int16_t *geno, const int32_t *nall, orc_state *s);

/* --- per-phase updates --------------------------------------------------- */
void orc_count_from_z(const orc_model *m, const int16_t *geno, const int32_t *Z,
                      int32_t *nA, int32_t *nQ);
void orc_update_z(const orc_model *m, const int16_t *geno, const double *Q, const double *P,
                  uint32_t rep, int32_t *Z /* may be NULL */, int32_t *nA, int32_t *nQ);
void orc_init_z(const orc_model *m, const int16_t *geno, int32_t *Z, int32_t *nA, int32_t *nQ);
void orc_update_p(const orc_model *m, const int32_t *nall, const int32_t *nA, const double *eps,
                  const double *F, const double *lambda, uint32_t rep, double *P);
void orc_update_q_gibbs(const orc_model *m, const int32_t *nQ, const double *alpha,
                        uint32_t rep, double *Q);
double orc_calc_like_ind_diff(const orc_model *m, const int16_t *geno, const double *Qplus,
                              const double *Qminus, const double *P, int ind);
void orc_update_q_metro(const orc_model *m, const int16_t *geno, const double *P,
                        const double *alpha, uint32_t rep, double *Q, int32_t *accepted);
double orc_log_prob_q(const orc_model *m, const double *Q, double onealpha);
double orc_log_prob_q_onepop(const orc_model *m, const double *Q, double alpha,
                             double alphasum, int pop);
double orc_alpha_prior_diff(const orc_model *m, double newalpha, double oldalpha);
void orc_update_alpha_given(const orc_model *m, const double *Q, double *alpha,
                            const double *z, const double *u);
void orc_update_alpha(const orc_model *m, const double *Q, uint32_t rep, double *alpha);
void orc_update_lambda_given(const orc_model *m, const int32_t *nall, const double *P,
                             const double *eps, double *lambda, const double *z,
                             const double *u);
void orc_update_lambda(const orc_model *m, const int32_t *nall, const double *P,
                       const double *eps, uint32_t rep, double *lambda);
void orc_update_eps_given(const orc_model *m, const int32_t *nall, const double *P,
                          const double *F, double lambda, double *eps,
                          const int32_t *a1, const int32_t *a2, const double *d,
                          const double *u);
void orc_update_eps(const orc_model *m, const int32_t *nall, const double *P, const double *F,
                    double lambda, uint32_t rep, double *eps);
double orc_fprior_diff(const orc_model *m, double newf, double oldf);
double orc_flike_freqs(const orc_model *m, const int32_t *nall, double f, const double *eps,
                       const double *P, int pop);
void orc_update_fst_given(const orc_model *m, const int32_t *nall, const double *eps,
                          const double *P, double *F, const double *z, const double *u);
void orc_update_fst(const orc_model *m, const int32_t *nall, const double *eps, const double *P,
                    uint32_t rep, double *F);
double orc_calc_like(const orc_model *m, const int16_t *geno, const double *Q, const double *P);
void orc_collect(const orc_model *m, const int16_t *geno, const int32_t *nall, orc_state *s);
void orc_sweep(const orc_model *m, const int16_t *geno, const int32_t *nall, uint32_t rep,
               orc_state *s);

/* --- 2-bit genotype words, host threads (oracle_packed.c) -----------------
 * words [ceil(N/16)][Lpad]: individual 16g+t at bits 2t..2t+1 of word (g, l),
 * code 0 = (0,0), 1 = (0,1), 2 = (1,1), 3 = missing.  nthreads <= 0: one per
 * online CPU (at most 64). */
double orc_calc_like_snp2(const orc_model *m, const uint32_t *words, int Lpad, const double *Q,
                          const double *P, int nthreads);
/* Non-missing copies per locus, allele-1 copies per locus, non-missing copies
 * per individual. */
void orc_snp2_counts(int N, int L, int Lpad, const uint32_t *words, int64_t *loc_valid,
                     int64_t *loc_a1, int64_t *ind_valid, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
