/*
 * CPU oracle for the STRUCTURE admixture sweep -- TEST INFRASTRUCTURE ONLY.
 *
 * An FP64 restatement of the reference CPU path (pmelsted/structure, C) with
 * the reference's array-fed / rand()-fed draws replaced by the Philox stream
 * described in oracle.h.  Each function names the reference lines it
 * follows.  Compiled with -ffp-contract=off so every product and sum is
 * rounded as the reference's gcc build rounds it (x86-64 SSE2, no FMA).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this file, as the checker; the product (structure_amd/) never links it.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_UNDERFLO 1e-100           /* structure.h:6 */
#define ORC_SQRTUNDER 1e-50           /* sqrt(UNDERFLO), CalcLike.c:99 */
#define ORC_LAMBDAMAX 10.0            /* structure.h:17 */
#define ORC_LAMBDAPROPSD 0.3          /* structure.h:18 */
#define ORC_GAMMA_MAX_BLOCKS 128u          /* 256 attempts */
#define ORC_GAMMA_BOOST_BLOCK 0xFFFFFFu

/* ------------------------------------------------------------------------
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11): two 32x32->64 multiplies per round, Weyl key
 * schedule.  Stands in for the reference's rand()/randomArr streams
 * (ran.c:128-161, randGen.c:23-36).
 * ---------------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * x0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
        uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        uint32_t y1 = (uint32_t)p1;
        uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        uint32_t y3 = (uint32_t)p0;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

void orc_philox_block(uint32_t seed, uint32_t chain, uint32_t c0, uint32_t c1, uint32_t c2,
                      uint32_t c3, uint32_t out[4])
{
    uint32_t ctr[4] = {c0, c1, c2, c3};
    uint32_t key[2] = {seed, chain};
    orc_philox4x32_10(ctr, key, out);
}

/* uniform in [0,1): the reference's randomArr values / RandomReal / RandomInteger (ran.c:128-147) */
double orc_uz(uint32_t w) { return (double)w * 0x1p-32; }
/* uniform in (0,1): the reference's rnd(), which rejects 0 and 1 (ran.c:152-161) */
double orc_uo(uint32_t w) { return ((double)w + 0.5) * 0x1p-32; }

/* Standard normal by Box-Muller (cos branch); stands in for snorm (ran.c:686-840). */
static double orc_box_muller(uint32_t w0, uint32_t w1)
{
    double r = sqrt(-2.0 * log(orc_uo(w0)));
    return r * cos(6.283185307179586 * orc_uo(w1));
}

double orc_normal(uint32_t seed, uint32_t chain, uint32_t c0, uint32_t c1, uint32_t c2,
                  uint32_t phase)
{
    uint32_t w[4];
    orc_philox_block(seed, chain, c0, c1, c2, phase << 24, w);
    return orc_box_muller(w[0], w[1]);
}

/* One Marsaglia-Tsang attempt with normal x and acceptance uniform u. */
static int orc_mt_attempt(double x, double u, double d, double c, double *g)
{
    double v = 1.0 + c * x;
    if (v <= 0.0) {
        return 0;
    }
    v = v * v * v;
    double x2 = x * x;
    if (u < 1.0 - 0.0331 * (x2 * x2) || log(u) < 0.5 * x2 + d * (1.0 - v + log(v))) {
        *g = d * v;
        return 1;
    }
    return 0;
}

/* Gamma(a, 1) by Marsaglia & Tsang (2000); shape < 1 via Gamma(a+1) * U^(1/a).
 * Replaces the reference samplers RGamma (ran.c:168-261) and RGammaDisc
 * (ran.c:1094-1188), which draw the same distribution by Ripley/Cheng GKM.
 * Philox block j serves two attempts: words 0,1 give a Box-Muller pair, the
 * cos normal is tried with the uniform of word 2, then the sin normal with the
 * uniform of word 3; the boost uniform for a < 1 is word 0 of block
 * ORC_GAMMA_BOOST_BLOCK. */
double orc_gamma(double a, uint32_t seed, uint32_t chain, uint32_t c0, uint32_t c1,
                 uint32_t c2, uint32_t phase)
{
    if (!(a > 0.0)) {
        return 0.0;
    }
    double aa = a < 1.0 ? a + 1.0 : a;
    double d = aa - 1.0 / 3.0;
    double c = 1.0 / sqrt(9.0 * d);
    double g = 0.0;
    int done = 0;
    for (uint32_t j = 0; j < ORC_GAMMA_MAX_BLOCKS && !done; j++) {
        uint32_t w[4];
        orc_philox_block(seed, chain, c0, c1, c2, (phase << 24) | j, w);
        double rad = sqrt(-2.0 * log(orc_uo(w[0])));
        double th = 6.283185307179586 * orc_uo(w[1]);
        done = orc_mt_attempt(rad * cos(th), orc_uo(w[2]), d, c, &g) ||
               orc_mt_attempt(rad * sin(th), orc_uo(w[3]), d, c, &g);
    }
    if (!done) {
        return d;   /* 256 rejections: unreachable in practice (acceptance >= 0.95) */
    }
    if (a < 1.0) {
        uint32_t b[4];
        orc_philox_block(seed, chain, c0, c1, c2, (phase << 24) | ORC_GAMMA_BOOST_BLOCK, b);
        g = g * exp(log(orc_uo(b[0])) / a);
    }
    return g;
}

/* ln Gamma(z), z > 0: Lanczos (g = 7, n = 9) with the published coefficient
 * table, summed in the same order as mylgamma (mymath.c:382-429). */
double orc_lgamma(double z)
{
    static const double c[9] = {
        0.9999999999995183, 676.5203681218835, -1259.139216722289,
        771.3234287757674, -176.6150291498386, 12.50734324009056,
        -0.1385710331296526, 9.934937113930748e-6, 1.659470187408462e-7};
    double s = 0.0;
    double den = z + 7.0;
    for (int j = 8; j >= 1; j--) {
        s += c[j] / den;
        den -= 1.0;
    }
    s += c[0];
    return log(s) + 0.9189385332046727 - (z + 6.5) + (z - 0.5) * log(z + 6.5);
}

/* GetNumFromPops's filter (UpdateP.c:85-88): with PFROMPOPFLAGONLY only the
 * individuals whose PopFlag is 1 add to the allele counts. */
static int orc_counts_alleles(const orc_model *m, int ind)
{
    return !m->pfrompopflagonly || (m->popflag && m->popflag[ind] == 1);
}

/* ------------------------------------------------------------------------
 * Counts from an explicit Z: GetNumFromPops (UpdateP.c:60-99, with its
 * PFROMPOPFLAGONLY filter) and GetNumLociPop (UpdateQ.c:17-35).
 * geno is [N][2][L] int16 with ORC_MISSING for missing copies.
 * ---------------------------------------------------------------------- */
void orc_count_from_z(const orc_model *m, const int16_t *geno, const int32_t *Z,
                      int32_t *nA, int32_t *nQ)
{
    const int N = m->N, L = m->L, K = m->K, A = m->A;
    memset(nA, 0, sizeof(int32_t) * (size_t)L * K * A);
    memset(nQ, 0, sizeof(int32_t) * (size_t)N * K);
    for (int i = 0; i < N; i++) {
        for (int c = 0; c < 2; c++) {
            for (int l = 0; l < L; l++) {
                size_t g = ((size_t)i * 2 + c) * L + l;
                int x = geno[g];
                int z = Z[g];
                if (x != ORC_MISSING && z != ORC_UNASSIGNED && orc_counts_alleles(m, i)) {
                    nA[((size_t)l * K + z) * A + x]++;
                }
                if (z != ORC_UNASSIGNED) {
                    nQ[(size_t)i * K + z]++;
                }
            }
        }
    }
}

/* UpdateZ (UpdateZ.c:13-56) with PickAnOptionDiscrete (mymath.c:105-129) and
 * numToRange (ran.c:1083-1087): Cutoffs[k] = Q[i,k]*P[l,k,x], summed in pop
 * order, r = 0 + u*(sum-0), first k with r <= running sum.  The uniform for
 * (i, l, copy) comes from phase 1.  Counts are accumulated directly. */
void orc_update_z(const orc_model *m, const int16_t *geno, const double *Q, const double *P,
                  uint32_t rep, int32_t *Z, int32_t *nA, int32_t *nQ)
{
    const int N = m->N, L = m->L, K = m->K, A = m->A;
    double cut[64];
    memset(nA, 0, sizeof(int32_t) * (size_t)L * K * A);
    memset(nQ, 0, sizeof(int32_t) * (size_t)N * K);
    for (int i = 0; i < N; i++) {
        for (int l = 0; l < L; l++) {
            uint32_t w[4];
            orc_philox_block(m->seed, m->chain, rep, (uint32_t)l, (uint32_t)(i >> 1),
                             (uint32_t)ORC_PH_Z << 24, w);
            int taken = 0;   /* uniforms are consumed per non-missing copy (randGen.c:23-36) */
            for (int c = 0; c < 2; c++) {
                size_t g = ((size_t)i * 2 + c) * L + l;
                int x = geno[g];
                int z;
                if (x == ORC_MISSING) {
                    z = ORC_UNASSIGNED;
                } else {
                    double sum = 0.0;
                    for (int k = 0; k < K; k++) {
                        cut[k] = Q[(size_t)i * K + k] * P[((size_t)l * K + k) * A + x];
                        sum += cut[k];
                    }
                    double r = 0.0 + orc_uz(w[2 * (i & 1) + taken++]) * (sum - 0.0);
                    double run = 0.0;
                    for (z = 0; z < K; z++) {
                        run += cut[z];
                        if (r <= run) {
                            break;
                        }
                    }
                    if (z >= K) {
                        z = K - 1;   /* reachable only with a NaN sum */
                    }
                    if (orc_counts_alleles(m, i)) {
                        nA[((size_t)l * K + z) * A + x]++;
                    }
                    nQ[(size_t)i * K + z]++;
                }
                if (Z) {
                    Z[g] = z;
                }
            }
        }
    }
}

/* InitializeZ (Init.c:124-169): every non-missing copy gets
 * RandomInteger(0, K-1) = (int)(u*K) (ran.c:138-147), u from phase 2. */
void orc_init_z(const orc_model *m, const int16_t *geno, int32_t *Z, int32_t *nA, int32_t *nQ)
{
    const int N = m->N, L = m->L, K = m->K, A = m->A;
    memset(nA, 0, sizeof(int32_t) * (size_t)L * K * A);
    memset(nQ, 0, sizeof(int32_t) * (size_t)N * K);
    for (int i = 0; i < N; i++) {
        for (int l = 0; l < L; l++) {
            uint32_t w[4];
            orc_philox_block(m->seed, m->chain, 0u, (uint32_t)l, (uint32_t)(i >> 1),
                             (uint32_t)ORC_PH_INITZ << 24, w);
            int taken = 0;
            for (int c = 0; c < 2; c++) {
                size_t g = ((size_t)i * 2 + c) * L + l;
                int x = geno[g];
                int z = ORC_UNASSIGNED;
                if (x != ORC_MISSING) {
                    z = (int)(orc_uz(w[2 * (i & 1) + taken++]) * (double)K);
                    if (orc_counts_alleles(m, i)) {
                        nA[((size_t)l * K + z) * A + x]++;
                    }
                    nQ[(size_t)i * K + z]++;
                }
                if (Z) {
                    Z[g] = z;
                }
            }
        }
    }
}

/* UpdateP (UpdateP.c:106-199): P[l,k,:] ~ Dirichlet(par), par[a] =
 * eps[l,a]*(1-F_k)/F_k + n[l,k,a] (FREQSCORR) or lambda_k + n[l,k,a];
 * the draw follows LogRDirichletDisc (ran.c:1247-1272) including the
 * sum < UNDERFLO fallback to uniform frequencies. */
void orc_update_p(const orc_model *m, const int32_t *nall, const int32_t *nA, const double *eps,
                  const double *F, const double *lambda, uint32_t rep, double *P)
{
    const int L = m->L, K = m->K, A = m->A;
    double g[256];
    for (int l = 0; l < L; l++) {
        int na = nall[l];
        for (int k = 0; k < K; k++) {
            double sum = 0.0;
            for (int a = 0; a < na; a++) {
                double par;
                if (m->freqscorr) {
                    par = eps[(size_t)l * A + a] * (1.0 - F[k]) / F[k] + nA[((size_t)l * K + k) * A + a];
                } else {
                    par = lambda[k] + nA[((size_t)l * K + k) * A + a];
                }
                g[a] = orc_gamma(par, m->seed, m->chain, rep, (uint32_t)l,
                                 ((uint32_t)k << 16) | (uint32_t)a, ORC_PH_P);
                sum += g[a];
            }
            double *p = P + ((size_t)l * K + k) * A;
            if (sum < ORC_UNDERFLO) {
                for (int a = 0; a < na; a++) {
                    p[a] = 1.0 / (double)na;
                }
            } else {
                for (int a = 0; a < na; a++) {
                    p[a] = g[a] / sum;
                }
            }
        }
    }
}

/* UpdateQAdmixture (UpdateQ.c:387-424) via RDirichlet (ran.c:348-359):
 * Q[i,:] ~ Dirichlet(alpha + NumLociPop[i,:]). */
void orc_update_q_gibbs(const orc_model *m, const int32_t *nQ, const double *alpha,
                        uint32_t rep, double *Q)
{
    const int N = m->N, K = m->K;
    for (int i = 0; i < N; i++) {
        double *q = Q + (size_t)i * K;
        double sum = 0.0;
        for (int k = 0; k < K; k++) {
            q[k] = orc_gamma(alpha[k] + nQ[(size_t)i * K + k], m->seed, m->chain, rep,
                             (uint32_t)i, (uint32_t)k, ORC_PH_Q);
            sum += q[k];
        }
        for (int k = 0; k < K; k++) {
            q[k] /= sum;
        }
    }
}

/* CalcLikeIndDiff (CalcLike.c:281-345): sum over copies of
 * log(Qplus.P) - log(Qminus.P), line-major. */
double orc_calc_like_ind_diff(const orc_model *m, const int16_t *geno, const double *Qplus,
                              const double *Qminus, const double *P, int ind)
{
    const int L = m->L, K = m->K, A = m->A;
    double logp = 0.0, logm = 0.0;
    for (int c = 0; c < 2; c++) {
        for (int l = 0; l < L; l++) {
            int x = geno[((size_t)ind * 2 + c) * L + l];
            if (x != ORC_MISSING) {
                double tp = 0.0, tm = 0.0;
                for (int k = 0; k < K; k++) {
                    double p = P[((size_t)l * K + k) * A + x];
                    tp += Qplus[k] * p;
                    tm += Qminus[k] * p;
                }
                logp += log(tp);
                logm += log(tm);
            }
        }
    }
    return logp - logm;
}

/* UpdateQMetro (UpdateQ.c:199-307): proposal TestQ ~ Dirichlet(alpha)
 * (RDirichletDisc, ran.c:1198-1210), accepted if u < exp(logdiff). */
void orc_update_q_metro(const orc_model *m, const int16_t *geno, const double *P,
                        const double *alpha, uint32_t rep, double *Q, int32_t *accepted)
{
    const int N = m->N, K = m->K;
    double tq[64];
    for (int i = 0; i < N; i++) {
        double sum = 0.0;
        for (int k = 0; k < K; k++) {
            tq[k] = orc_gamma(alpha[k], m->seed, m->chain, rep, (uint32_t)i, (uint32_t)k,
                              ORC_PH_QMETRO);
            sum += tq[k];
        }
        for (int k = 0; k < K; k++) {
            tq[k] /= sum;
        }
        double *q = Q + (size_t)i * K;
        double logdiff = orc_calc_like_ind_diff(m, geno, tq, q, P, i);
        uint32_t w[4];
        orc_philox_block(m->seed, m->chain, rep, (uint32_t)i, 0xffffffffu,
                         (uint32_t)ORC_PH_QMETRO << 24, w);
        int acc = orc_uo(w[0]) < exp(logdiff);
        if (acc) {
            for (int k = 0; k < K; k++) {
                q[k] = tq[k];
            }
        }
        if (accepted) {
            accepted[i] = acc;
        }
    }
}

/* LogProbQ (UpdateAlpha.c:144-186): clamped running product of Q with a
 * log taken whenever it drops below sqrt(UNDERFLO). USEPOPINFO=0. */
double orc_log_prob_q(const orc_model *m, const double *Q, double onealpha)
{
    const int N = m->N, K = m->K;
    double sum = 0.0, run = 1.0;
    for (int i = 0; i < N; i++) {
        for (int k = 0; k < K; k++) {
            double q = Q[(size_t)i * K + k];
            run *= (q > ORC_SQRTUNDER) ? q : ORC_SQRTUNDER;
            if (run < ORC_SQRTUNDER) {
                sum += (onealpha - 1.0) * log(run);
                run = 1.0;
            }
        }
    }
    sum += (onealpha - 1.0) * log(run);
    sum += (orc_lgamma(K * onealpha) - K * orc_lgamma(onealpha)) * N;
    return sum;
}

/* LogProbQonepop (UpdateAlpha.c:84-116). */
double orc_log_prob_q_onepop(const orc_model *m, const double *Q, double alpha,
                             double alphasum, int pop)
{
    const int N = m->N, K = m->K;
    double logterm = 0.0;
    for (int i = 0; i < N; i++) {
        logterm += log(Q[(size_t)i * K + pop]);
    }
    double sum = 0.0;
    sum += (alpha - 1.0) * logterm;
    sum += (orc_lgamma(alphasum) - orc_lgamma(alpha)) * N;
    return sum;
}

/* AlphaPriorDiff (UpdateAlpha.c:12-18): gamma(A, B) prior log ratio. */
double orc_alpha_prior_diff(const orc_model *m, double newalpha, double oldalpha)
{
    return (m->alphapriora - 1) * log(newalpha / oldalpha) +
           (oldalpha - newalpha) / m->alphapriorb;
}

/* UpdateAlpha (UpdateAlpha.c:189-253) given the standard normals z[pop] and
 * acceptance uniforms u[pop].  As in the reference, logprobdiff starts at 0
 * once and is only reset by the gamma prior term. */
void orc_update_alpha_given(const orc_model *m, const double *Q, double *alpha,
                            const double *z, const double *u)
{
    const int K = m->K;
    int numalphas = m->popalphas ? K : 1;
    double logprobdiff = 0.0;
    for (int pop = 0; pop < numalphas; pop++) {
        double newalpha = alpha[pop] + m->alphapropsd * z[pop];
        if ((newalpha > 0) && ((newalpha < m->alphamax) || (!m->unifprioralpha))) {
            if (!m->unifprioralpha) {
                logprobdiff = orc_alpha_prior_diff(m, newalpha, alpha[pop]);
            }
            if (m->popalphas) {
                double sumalphas = 0.0;
                for (int i = 0; i < K; i++) {
                    sumalphas += alpha[i];
                }
                logprobdiff -= orc_log_prob_q_onepop(m, Q, alpha[pop], sumalphas, pop);
                sumalphas += newalpha - alpha[pop];
                logprobdiff += orc_log_prob_q_onepop(m, Q, newalpha, sumalphas, pop);
            } else {
                logprobdiff += orc_log_prob_q(m, Q, newalpha);
                logprobdiff -= orc_log_prob_q(m, Q, alpha[pop]);
            }
            if (u[pop] < exp(logprobdiff)) {
                alpha[pop] = newalpha;
                if (!m->popalphas) {
                    for (int p = 1; p < K; p++) {
                        alpha[p] = newalpha;
                    }
                }
            }
        }
    }
}

void orc_update_alpha(const orc_model *m, const double *Q, uint32_t rep, double *alpha)
{
    double z[64], u[64];
    int numalphas = m->popalphas ? m->K : 1;
    for (int pop = 0; pop < numalphas; pop++) {
        uint32_t w[4];
        orc_philox_block(m->seed, m->chain, rep, (uint32_t)pop, 0u,
                         (uint32_t)ORC_PH_ALPHA << 24, w);
        z[pop] = orc_box_muller(w[0], w[1]);
        u[pop] = orc_uo(w[2]);
    }
    orc_update_alpha_given(m, Q, alpha, z, u);
}

/* UpdateLambda (UpdateLambda.c:49-102) and UpdatePopLambda (:8-45), given
 * normals z[pop] and uniforms u[pop] (index 0 for the single lambda). */
void orc_update_lambda_given(const orc_model *m, const int32_t *nall, const double *P,
                             const double *eps, double *lambda, const double *z,
                             const double *u)
{
    const int L = m->L, K = m->K, A = m->A;
    if (m->popspecificlambda) {
        for (int pop = 0; pop < K; pop++) {
            double nw = lambda[pop] + ORC_LAMBDAPROPSD * z[pop];
            if ((nw > 0.0) && (nw < ORC_LAMBDAMAX)) {
                double sum = 0.0;
                for (int l = 0; l < L; l++) {
                    int na = nall[l];
                    if (na > 1) {
                        sum += orc_lgamma((double)na * nw);
                        sum -= orc_lgamma((double)na * lambda[pop]);
                        sum += (double)na * orc_lgamma(lambda[pop]);
                        sum -= (double)na * orc_lgamma(nw);
                        double sumlogp = 0.0;
                        for (int a = 0; a < na; a++) {
                            sumlogp += log(P[((size_t)l * K + pop) * A + a]);
                        }
                        sum += (nw - lambda[pop]) * sumlogp;
                    }
                }
                if (u[pop] < exp(sum)) {
                    lambda[pop] = nw;
                }
            }
        }
        return;
    }
    double nw = lambda[0] + ORC_LAMBDAPROPSD * z[0];
    if ((nw > 0.0) && (nw < ORC_LAMBDAMAX)) {
        int stoppop = m->freqscorr ? 1 : K;
        double sum = 0.0;
        for (int l = 0; l < L; l++) {
            int na = nall[l];
            if (na > 1) {
                sum += (double)stoppop * orc_lgamma((double)na * nw);
                sum -= (double)stoppop * orc_lgamma((double)na * lambda[0]);
                sum += (double)stoppop * (double)na * orc_lgamma(lambda[0]);
                sum -= (double)stoppop * (double)na * orc_lgamma(nw);
                double sumlogp = 0.0;
                for (int pop = 0; pop < stoppop; pop++) {
                    for (int a = 0; a < na; a++) {
                        if (m->freqscorr) {
                            sumlogp += log(eps[(size_t)l * A + a]);
                        } else {
                            sumlogp += log(P[((size_t)l * K + pop) * A + a]);
                        }
                    }
                }
                sum += (nw - lambda[0]) * sumlogp;
            }
        }
        if (u[0] < exp(sum)) {
            for (int pop = 0; pop < K; pop++) {
                lambda[pop] = nw;
            }
        }
    }
}

void orc_update_lambda(const orc_model *m, const int32_t *nall, const double *P,
                       const double *eps, uint32_t rep, double *lambda)
{
    double z[64], u[64];
    int n = m->popspecificlambda ? m->K : 1;
    for (int pop = 0; pop < n; pop++) {
        uint32_t w[4];
        orc_philox_block(m->seed, m->chain, rep, (uint32_t)pop, 0u,
                         (uint32_t)ORC_PH_LAMBDA << 24, w);
        z[pop] = orc_box_muller(w[0], w[1]);
        u[pop] = orc_uo(w[2]);
    }
    orc_update_lambda_given(m, nall, P, eps, lambda, z, u);
}

/* NonIndependenceUpdateEpsilon (UpdateEpsilon.c:71-145), given for each locus
 * the two alleles (a2 already shifted past a1), the step d and the uniform. */
void orc_update_eps_given(const orc_model *m, const int32_t *nall, const double *P,
                          const double *F, double lambda, double *eps,
                          const int32_t *a1v, const int32_t *a2v, const double *dv,
                          const double *uv)
{
    const int L = m->L, K = m->K, A = m->A;
    for (int l = 0; l < L; l++) {
        if (nall[l] > 1) {
            int a1 = a1v[l], a2 = a2v[l];
            double d = dv[l];
            double *e = eps + (size_t)l * A;
            if ((e[a1] + d < 1.0) && (e[a2] - d > 0.0)) {
                double sum = 0.0;
                for (int pop = 0; pop < K; pop++) {
                    double frac = (1.0 - F[pop]) / F[pop];
                    sum += orc_lgamma(frac * e[a1]);
                    sum += orc_lgamma(frac * e[a2]);
                    sum -= orc_lgamma(frac * (e[a1] + d));
                    sum -= orc_lgamma(frac * (e[a2] - d));
                    sum += frac * d * log(P[((size_t)l * K + pop) * A + a1]);
                    sum -= frac * d * log(P[((size_t)l * K + pop) * A + a2]);
                }
                if (lambda != 1.0) {
                    sum += log(pow((e[a1] + d) * (e[a2] - d) / (e[a1]) / (e[a2]),
                                   (double)lambda - 1.0));
                }
                if (uv[l] < exp(sum)) {
                    e[a1] += d;
                    e[a2] -= d;
                }
            }
        }
    }
}

void orc_update_eps(const orc_model *m, const int32_t *nall, const double *P, const double *F,
                    double lambda, uint32_t rep, double *eps)
{
    const int L = m->L;
    int32_t *a1 = (int32_t *)calloc((size_t)L, sizeof(int32_t));
    int32_t *a2 = (int32_t *)calloc((size_t)L, sizeof(int32_t));
    double *d = (double *)calloc((size_t)L, sizeof(double));
    double *u = (double *)calloc((size_t)L, sizeof(double));
    double invsqrtn = pow((double)m->N, -0.5);
    for (int l = 0; l < L; l++) {
        int na = nall[l];
        if (na > 1) {
            uint32_t w[4];
            orc_philox_block(m->seed, m->chain, rep, (uint32_t)l, 0u,
                             (uint32_t)ORC_PH_EPS << 24, w);
            a1[l] = (int)(orc_uz(w[0]) * (double)na);
            a2[l] = (int)(orc_uz(w[1]) * (double)(na - 1));
            if (a2[l] >= a1[l]) {
                a2[l] += 1;
            }
            d[l] = 0.0 + orc_uz(w[2]) * (invsqrtn - 0.0);
            u[l] = orc_uo(w[3]);
        }
    }
    orc_update_eps_given(m, nall, P, F, lambda, eps, a1, a2, d, u);
    free(a1); free(a2); free(d); free(u);
}

/* FPriorDiff (UpdateFst.c:10-18). */
double orc_fprior_diff(const orc_model *m, double newf, double oldf)
{
    double mu = m->fpriormean, sd = m->fpriorsd;
    return ((mu * mu / (sd * sd) - 1) * log(newf / oldf) + (oldf - newf) * mu / (sd * sd));
}

/* FlikeFreqs (UpdateFst.c:22-51). */
double orc_flike_freqs(const orc_model *m, const int32_t *nall, double f, const double *eps,
                       const double *P, int pop)
{
    const int L = m->L, K = m->K, A = m->A;
    double frac = (1.0 - f) / f;
    double sum = L * orc_lgamma(frac);
    for (int l = 0; l < L; l++) {
        for (int a = 0; a < nall[l]; a++) {
            sum += frac * eps[(size_t)l * A + a] * log(P[((size_t)l * K + pop) * A + a]);
            sum -= orc_lgamma(frac * eps[(size_t)l * A + a]);
        }
        if (nall[l] == 0) {
            sum -= orc_lgamma(frac);
        }
    }
    return sum;
}

/* UpdateFst (UpdateFst.c:142-204) given normals z[pop1] and uniforms u[pop1]. */
void orc_update_fst_given(const orc_model *m, const int32_t *nall, const double *eps,
                          const double *P, double *F, const double *z, const double *u)
{
    const int K = m->K;
    int numpops1 = m->onefst ? 1 : K;
    for (int pop1 = 0; pop1 < numpops1; pop1++) {
        double oldf = F[pop1];
        double newf = oldf + m->fpriorsd * z[pop1];
        if (newf > 0.0 && newf < 1.0) {
            double logprobdiff = orc_fprior_diff(m, newf, oldf);
            int numpops2 = m->onefst ? K : pop1 + 1;
            for (int pop2 = pop1; pop2 < numpops2; pop2++) {
                logprobdiff += orc_flike_freqs(m, nall, newf, eps, P, pop2);
                logprobdiff -= orc_flike_freqs(m, nall, oldf, eps, P, pop2);
            }
            int acc = (logprobdiff >= 0.0) || (u[pop1] < exp(logprobdiff));
            if (acc) {
                for (int pop2 = pop1; pop2 < numpops2; pop2++) {
                    F[pop2] = newf;
                }
            }
        }
    }
}

void orc_update_fst(const orc_model *m, const int32_t *nall, const double *eps, const double *P,
                    uint32_t rep, double *F)
{
    double z[64], u[64];
    int n = m->onefst ? 1 : m->K;
    for (int pop = 0; pop < n; pop++) {
        uint32_t w[4];
        orc_philox_block(m->seed, m->chain, rep, (uint32_t)pop, 0u,
                         (uint32_t)ORC_PH_FST << 24, w);
        z[pop] = orc_box_muller(w[0], w[1]);
        u[pop] = orc_uo(w[2]);
    }
    orc_update_fst_given(m, nall, eps, P, F, z, u);
}

/* CalcLike (CalcLike.c:387-426) -> CalcLikeInd (:75-133): per individual a
 * clamped running product of Q.P over (line, locus), logged whenever it
 * drops below sqrt(UNDERFLO). */
double orc_calc_like(const orc_model *m, const int16_t *geno, const double *Q, const double *P)
{
    const int N = m->N, L = m->L, K = m->K, A = m->A;
    double loglike = 0.0;
    for (int i = 0; i < N; i++) {
        double run = 1.0, ind = 0.0;
        for (int c = 0; c < 2; c++) {
            for (int l = 0; l < L; l++) {
                int x = geno[((size_t)i * 2 + c) * L + l];
                if (x != ORC_MISSING) {
                    double term = 0.0;
                    for (int k = 0; k < K; k++) {
                        term += Q[(size_t)i * K + k] * P[((size_t)l * K + k) * A + x];
                    }
                    run *= (term > ORC_SQRTUNDER) ? term : ORC_SQRTUNDER;
                    if (run < ORC_SQRTUNDER) {
                        ind += log(run);
                        run = 1.0;
                    }
                }
            }
        }
        ind += log(run);
        loglike += ind;
    }
    return loglike;
}

/* DataCollection (output.c:50-101) -> UpdateSums (output.c:654-707):
 * posterior sums, the ANCESTDIST histogram (output.c:691-700) and the lnP(D)
 * accumulators (LOCPRIOR=0). */
void orc_collect(const orc_model *m, const int16_t *geno, const int32_t *nall, orc_state *s)
{
    const int N = m->N, L = m->L, K = m->K, A = m->A;
    for (int k = 0; k < K; k++) {
        s->lambdaSum[k] += s->lambda[k];
    }
    for (size_t j = 0; j < (size_t)N * K; j++) {
        s->QSum[j] += s->Q[j];
    }
    for (int l = 0; l < L; l++) {
        for (int k = 0; k < K; k++) {
            for (int a = 0; a < nall[l]; a++) {
                size_t j = ((size_t)l * K + k) * A + a;
                s->PSum[j] += s->P[j];
            }
        }
    }
    if (m->freqscorr) {
        for (int k = 0; k < K; k++) {
            s->FstSum[k] += s->F[k];
        }
        for (int l = 0; l < L; l++) {
            for (int a = 0; a < nall[l]; a++) {
                s->epsSum[(size_t)l * A + a] += s->eps[(size_t)l * A + a];
            }
        }
    }
    if (m->ancestdist && s->ancestDist) {
        const int nb = m->numboxes;
        for (int i = 0; i < N; i++) {
            for (int k = 0; k < K; k++) {
                int box = (int)(s->Q[(size_t)i * K + k] * (double)nb);
                if (box == nb) {
                    box = nb - 1;   /* Q = 1.000 */
                }
                s->ancestDist[((size_t)i * K + k) * nb + box]++;
            }
        }
    }
    for (int k = 0; k < K; k++) {
        s->alphaSum[k] += s->alpha[k];
    }
    if (m->computeprob) {
        s->like = orc_calc_like(m, geno, s->Q, s->P);
        s->sumlikes += s->like;
        s->sumsqlikes += s->like * s->like;
    }
    s->nsum++;
}

/* Initialization (Init.c:329-444) for the admixture model: alpha, lambda,
 * InitFreqPriors (Init.c:172-240), Q = 1/K (structure.c:51-58), random Z
 * (Init.c:124-169) reduced to its counts, zeroed sums. */
void orc_init_state(const orc_model *m, const int16_t *geno, const int32_t *nall, orc_state *s)
{
    const int N = m->N, L = m->L, K = m->K, A = m->A;
    for (int k = 0; k < K; k++) {
        s->alpha[k] = m->alpha0;
        s->lambda[k] = m->lambda0;
        s->alphaSum[k] = 0.0;
        s->lambdaSum[k] = 0.0;
        s->FstSum[k] = 0.0;
    }
    memset(s->eps, 0, sizeof(double) * (size_t)L * A);
    if (!m->freqscorr) {
        for (int l = 0; l < L; l++) {
            for (int a = 0; a < nall[l]; a++) {
                s->eps[(size_t)l * A + a] = m->lambda0;
            }
        }
        for (int k = 0; k < K; k++) {
            s->F[k] = 0.5;
        }
    } else {
        int32_t cnt[256];
        for (int l = 0; l < L; l++) {
            int na = nall[l];
            for (int a = 0; a < na; a++) {
                cnt[a] = 0;
            }
            int total = 0;
            for (int i = 0; i < N; i++) {
                for (int c = 0; c < 2; c++) {
                    int x = geno[((size_t)i * 2 + c) * L + l];
                    if (x != ORC_MISSING) {
                        total++;
                        if (x >= 0 && x < na) {
                            cnt[x]++;
                        }
                    }
                }
            }
            for (int a = 0; a < na; a++) {
                s->eps[(size_t)l * A + a] = (((double)m->lambda0 + (double)cnt[a]) /
                                             ((double)na * (double)m->lambda0 + (double)total));
            }
        }
        for (int k = 0; k < K; k++) {
            s->F[k] = m->fpriormean;
        }
    }
    for (size_t j = 0; j < (size_t)N * K; j++) {
        s->Q[j] = 1.0 / K;
        s->QSum[j] = 0.0;
    }
    memset(s->P, 0, sizeof(double) * (size_t)L * K * A);
    memset(s->PSum, 0, sizeof(double) * (size_t)L * K * A);
    memset(s->epsSum, 0, sizeof(double) * (size_t)L * A);
    orc_init_z(m, geno, NULL, s->nA, s->nQ);
    if (m->ancestdist && s->ancestDist) {   /* InitializeSums (Init.c:265-271) */
        memset(s->ancestDist, 0, sizeof(int32_t) * (size_t)N * K * m->numboxes);
    }
    s->like = s->sumlikes = s->sumsqlikes = 0.0;
    s->nsum = 0;
}

/* One sweep in the order of the reference hot loop (structure.c:497-638):
 * P, Q (Metropolis when rep % METROFREQ == 0, UpdateQ.c:740), Z, alpha,
 * lambda, epsilon + Fst, then data collection once rep + 1 > BURNIN. */
void orc_sweep(const orc_model *m, const int16_t *geno, const int32_t *nall, uint32_t rep,
               orc_state *s)
{
    orc_update_p(m, nall, s->nA, s->eps, s->F, s->lambda, rep, s->P);
    if (m->metrofreq > 0 && rep % (uint32_t)m->metrofreq == 0) {
        orc_update_q_metro(m, geno, s->P, s->alpha, rep, s->Q, NULL);
    } else {
        orc_update_q_gibbs(m, s->nQ, s->alpha, rep, s->Q);
    }
    orc_update_z(m, geno, s->Q, s->P, rep, NULL, s->nA, s->nQ);
    if (m->inferalpha) {
        orc_update_alpha(m, s->Q, rep, s->alpha);
    }
    if (m->inferlambda) {
        orc_update_lambda(m, nall, s->P, s->eps, rep, s->lambda);
    }
    if (m->freqscorr) {
        orc_update_eps(m, nall, s->P, s->F, s->lambda[0], rep, s->eps);
        orc_update_fst(m, nall, s->eps, s->P, rep, s->F);
    }
    if ((int64_t)rep + 1 > m->burnin) {
        orc_collect(m, geno, nall, s);
    }
}
