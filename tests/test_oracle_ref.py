"""Pin the CPU oracle against the reference's own C code (oracle/_ref).

Every test here builds the same inputs for both, feeds the reference the
uniforms the oracle would draw (through randomArr or the rand() queue of
oracle/ref_harness.c) and requires identical results: bit-exact for counts,
Z and the FP64 log-ratio helpers, which the oracle restates in the
reference's own operation order.
"""
import math

import numpy as np
import pytest

import oracle
from helpers import RAND_MAX, geno_ref_layout, random_geno, random_p, random_q, z_uniforms

TWO31 = 2.0 ** 31


def _rand_queue(rng, n):
    # values in [1, RAND_MAX] so that rnd() (ran.c:152-161) never rejects one
    return rng.integers(1, RAND_MAX + 1, size=n).astype(np.int32)


@pytest.mark.parametrize("K,A,miss", [(1, 2, 0.0), (2, 2, 0.03), (3, 7, 0.05), (6, 4, 0.2)])
def test_update_z_and_counts_match_reference(ref, K, A, miss):
    rng = np.random.default_rng(100 + K)
    N, L = 23, 31
    geno, nall = random_geno(rng, N, L, A, miss=miss)
    Q = random_q(rng, N, K, a=0.7)
    P = random_p(rng, L, K, A, nall, a=0.8)
    rep, seed, chain = 7, 12345, 3
    ref.configure(N, L, K, A)
    g32 = geno_ref_layout(geno)
    Zr = np.zeros((N, 2, L), np.int32)
    ra = z_uniforms(seed, chain, rep, N, L)
    ref.lib().UpdateZ(ref.ip(Zr), ref.dp(Q), ref.dp(P), ref.ip(g32), ref.dp(ra))
    nAr = np.zeros((L, K, A), np.int32)
    ref.lib().GetNumFromPops(ref.ip(nAr), ref.ip(g32), ref.ip(Zr), ref.ip(nall),
                             ref.individuals(N))
    nQr = np.zeros((N, K), np.int32)
    for i in range(N):
        row = np.zeros(K, np.int32)
        ref.lib().GetNumLociPop(ref.ip(row), ref.ip(Zr), i)
        nQr[i] = row

    m = oracle.make_model(N, L, K, A, seed=seed, chain=chain)
    ch = oracle.Chain(m, geno, nall)
    nA, nQ, Z = ch.update_z(Q, P, rep, want_z=True)
    assert np.array_equal(Z, Zr)
    assert np.array_equal(nA, nAr)
    assert np.array_equal(nQ, nQr)
    # count_from_z restates GetNumFromPops/GetNumLociPop on an explicit Z
    nA2 = np.zeros_like(nA)
    nQ2 = np.zeros_like(nQ)
    oracle.lib().orc_count_from_z(oracle.C.byref(m), oracle.bp(ch.geno), oracle.ip(Zr),
                                  oracle.ip(nA2), oracle.ip(nQ2))
    assert np.array_equal(nA2, nAr) and np.array_equal(nQ2, nQr)


def test_calc_like_and_ind_diff_match_reference(ref):
    rng = np.random.default_rng(5)
    N, L, K, A = 17, 40, 4, 6
    geno, nall = random_geno(rng, N, L, A, miss=0.1)
    Q = random_q(rng, N, K, a=0.5)
    P = random_p(rng, L, K, A, nall, a=0.5)
    ref.configure(N, L, K, A)
    g32 = geno_ref_layout(geno)
    m = oracle.make_model(N, L, K, A)
    ch = oracle.Chain(m, geno, nall)
    like_ref = ref.lib().CalcLike(ref.ip(g32), None, ref.dp(Q), ref.dp(P), None, None, None)
    assert ch.calc_like(Q, P) == like_ref
    Q2 = random_q(rng, N, K, a=0.5)
    for i in range(N):
        qp = np.ascontiguousarray(Q2[i])
        qm = np.ascontiguousarray(Q[i])
        r = ref.lib().CalcLikeIndDiff(ref.ip(g32), None, ref.dp(qp), ref.dp(qm), ref.dp(P), i,
                                      None)
        o = oracle.lib().orc_calc_like_ind_diff(oracle.C.byref(m), oracle.bp(ch.geno),
                                                oracle.dp(qp), oracle.dp(qm), oracle.dp(P), i)
        assert o == r


def test_calc_like_underflow_clamp_matches_reference(ref):
    # many tiny terms: exercises the sqrt(UNDERFLO) clamp (CalcLike.c:113-126)
    rng = np.random.default_rng(9)
    N, L, K, A = 4, 300, 2, 2
    geno, nall = random_geno(rng, N, L, A, miss=0.0, biallelic=True)
    Q = np.tile([1.0 - 1e-30, 1e-30], (N, 1))
    P = np.zeros((L, K, A))
    P[:, 0, :] = [1e-60, 1.0 - 1e-60]
    P[:, 1, :] = [0.5, 0.5]
    ref.configure(N, L, K, A)
    ch = oracle.Chain(oracle.make_model(N, L, K, A), geno, nall)
    like_ref = ref.lib().CalcLike(ref.ip(geno_ref_layout(geno)), None, ref.dp(Q), ref.dp(P),
                                  None, None, None)
    assert ch.calc_like(Q, P) == like_ref


def test_lgamma_bitexact_and_accurate(ref):
    rng = np.random.default_rng(1)
    zs = np.concatenate([rng.uniform(1e-4, 1, 300), rng.uniform(1, 50, 300),
                         10 ** rng.uniform(1, 6, 300), [0.5, 1.0, 2.0, 3.0]])
    for z in zs:
        assert oracle.lgamma(z) == ref.lib().mylgamma(float(z))
        # against the library routine: the Lanczos table is good to ~1e-13
        assert abs(oracle.lgamma(z) - math.lgamma(z)) <= 1e-12 * max(1.0, abs(math.lgamma(z)))


@pytest.mark.parametrize("popalphas,unif", [(0, 1), (1, 1), (0, 0), (1, 0)])
def test_alpha_log_ratio_helpers_match_reference(ref, popalphas, unif):
    rng = np.random.default_rng(11 + popalphas)
    N, K = 31, 5
    ref.configure(N, 3, K, 2, POPALPHAS=popalphas, UNIFPRIORALPHA=unif, ALPHAPRIORA=1.5,
                  ALPHAPRIORB=3.0)
    m = oracle.make_model(N, 3, K, 2, popalphas=popalphas, unifprioralpha=unif,
                          alphapriora=1.5, alphapriorb=3.0)
    inds = ref.individuals(N)
    for _ in range(20):
        Q = random_q(rng, N, K, a=rng.uniform(0.05, 3))
        Q[rng.random(Q.shape) < 0.02] = 1e-70   # exercise the clamp in LogProbQ
        a = rng.uniform(0.01, 5)
        assert oracle.lib().orc_log_prob_q(oracle.C.byref(m), oracle.dp(Q), a) == \
            ref.lib().LogProbQ(ref.dp(Q), a, inds)
        for pop in range(K):
            s = rng.uniform(1, 10)
            assert oracle.lib().orc_log_prob_q_onepop(oracle.C.byref(m), oracle.dp(Q), a, s,
                                                      pop) == \
                ref.lib().LogProbQonepop(ref.dp(Q), a, s, inds, pop)
        b = rng.uniform(0.01, 5)
        assert oracle.lib().orc_alpha_prior_diff(oracle.C.byref(m), a, b) == \
            ref.lib().AlphaPriorDiff(a, b)


@pytest.mark.parametrize("popalphas,unif", [(0, 1), (1, 1), (0, 0), (1, 0)])
def test_update_alpha_matches_reference(ref, popalphas, unif):
    rng = np.random.default_rng(21 + 2 * popalphas + unif)
    N, K = 25, 4
    kw = dict(POPALPHAS=popalphas, UNIFPRIORALPHA=unif, ALPHAPROPSD=0.3, ALPHAMAX=3.0,
              ALPHAPRIORA=2.0, ALPHAPRIORB=1.5)
    ref.configure(N, 3, K, 2, **kw)
    m = oracle.make_model(N, 3, K, 2, popalphas=popalphas, unifprioralpha=unif,
                          alphapropsd=0.3, alphamax=3.0, alphapriora=2.0, alphapriorb=1.5)
    nacc = 0
    for trial in range(60):
        Q = random_q(rng, N, K, a=rng.uniform(0.3, 3))
        alpha = rng.uniform(0.05, 3.0, K) if popalphas else np.full(K, rng.uniform(0.05, 3))
        queue = _rand_queue(rng, 400)
        nal = K if popalphas else 1
        z, u, pos = np.zeros(K), np.zeros(K), 0
        for p in range(nal):
            ref.set_rand(queue[pos:])
            z[p] = ref.lib().snorm()
            pos += ref.lib().sref_rand_consumed()
            new = alpha[p] + 0.3 * z[p]
            if new > 0 and (new < 3.0 or not unif):
                u[p] = queue[pos] / TWO31
                pos += 1
        ar = alpha.copy()
        ref.set_rand(queue)
        ref.lib().UpdateAlpha(ref.dp(Q), ref.dp(ar), ref.individuals(N), 1)
        assert ref.lib().sref_rand_consumed() == pos
        ao = alpha.copy()
        oracle.lib().orc_update_alpha_given(oracle.C.byref(m), oracle.dp(Q), oracle.dp(ao),
                                            oracle.dp(z), oracle.dp(u))
        assert np.array_equal(ao, ar)
        nacc += int(not np.array_equal(ao, alpha))
    assert 5 < nacc < 60   # both outcomes exercised


@pytest.mark.parametrize("freqscorr,popspec", [(0, 0), (1, 0), (0, 1)])
def test_update_lambda_matches_reference(ref, freqscorr, popspec):
    rng = np.random.default_rng(31 + freqscorr + 2 * popspec)
    N, L, K, A = 10, 30, 3, 5
    _, nall = random_geno(rng, N, L, A)
    ref.configure(N, L, K, A, FREQSCORR=freqscorr, POPSPECIFICLAMBDA=popspec)
    m = oracle.make_model(N, L, K, A, freqscorr=freqscorr, popspecificlambda=popspec)
    nacc = 0
    for trial in range(60):
        P = random_p(rng, L, K, A, nall, a=rng.uniform(0.3, 3))
        eps = np.zeros((L, A))
        for l in range(L):
            eps[l, :nall[l]] = rng.dirichlet(np.ones(nall[l]))
        lam = np.full(K, rng.uniform(0.05, 3)) if not popspec else rng.uniform(0.05, 3, K)
        queue = _rand_queue(rng, 400)
        n = K if popspec else 1
        z, u, pos = np.zeros(K), np.zeros(K), 0
        for p in range(n):
            ref.set_rand(queue[pos:])
            z[p] = ref.lib().snorm()
            pos += ref.lib().sref_rand_consumed()
            new = lam[p] + 0.3 * z[p]
            if 0 < new < 10:
                u[p] = queue[pos] / TWO31
                pos += 1
        lr = lam.copy()
        ref.set_rand(queue)
        if popspec:
            ref.lib().UpdatePopLambda(ref.dp(P), ref.dp(lr), ref.ip(nall))
        else:
            ref.lib().UpdateLambda(ref.dp(P), ref.dp(eps), ref.dp(lr), ref.ip(nall))
        assert ref.lib().sref_rand_consumed() == pos
        lo = lam.copy()
        oracle.lib().orc_update_lambda_given(oracle.C.byref(m), oracle.ip(nall), oracle.dp(P),
                                             oracle.dp(eps), oracle.dp(lo), oracle.dp(z),
                                             oracle.dp(u))
        assert np.array_equal(lo, lr)
        nacc += int(not np.array_equal(lo, lam))
    assert 3 < nacc < 60


@pytest.mark.parametrize("lam0", [1.0, 0.7])
def test_update_eps_matches_reference(ref, lam0):
    rng = np.random.default_rng(41)
    N, L, K, A = 12, 40, 3, 5
    _, nall = random_geno(rng, N, L, A)
    ref.configure(N, L, K, A)
    m = oracle.make_model(N, L, K, A)
    invs = N ** -0.5
    for trial in range(10):
        P = random_p(rng, L, K, A, nall, a=1.0)
        F = rng.uniform(0.01, 0.3, K)
        eps = np.zeros((L, A))
        for l in range(L):
            eps[l, :nall[l]] = rng.dirichlet(np.ones(nall[l]) * 0.5)
        queue = _rand_queue(rng, 5 * L)
        a1 = np.zeros(L, np.int32)
        a2 = np.zeros(L, np.int32)
        d = np.zeros(L)
        u = np.zeros(L)
        pos = 0
        for l in range(L):
            na = nall[l]
            if na > 1:
                a1[l] = int(queue[pos] / TWO31 * na)
                a2[l] = int(queue[pos + 1] / TWO31 * (na - 1))
                if a2[l] >= a1[l]:
                    a2[l] += 1
                d[l] = 0.0 + (queue[pos + 2] / TWO31) * (invs - 0.0)
                pos += 3
                if eps[l, a1[l]] + d[l] < 1.0 and eps[l, a2[l]] - d[l] > 0.0:
                    u[l] = queue[pos] / TWO31
                    pos += 1
        er = eps.copy()
        ref.set_rand(queue)
        ref.lib().NonIndependenceUpdateEpsilon(ref.dp(P), ref.dp(er), ref.dp(F), ref.ip(nall),
                                               lam0)
        assert ref.lib().sref_rand_consumed() == pos
        eo = eps.copy()
        oracle.lib().orc_update_eps_given(oracle.C.byref(m), oracle.ip(nall), oracle.dp(P),
                                          oracle.dp(F), lam0, oracle.dp(eo), oracle.ip(a1),
                                          oracle.ip(a2), oracle.dp(d), oracle.dp(u))
        assert np.array_equal(eo, er)
        assert not np.array_equal(eo, eps)


def test_fst_helpers_match_reference(ref):
    rng = np.random.default_rng(51)
    N, L, K, A = 12, 50, 3, 4
    _, nall = random_geno(rng, N, L, A)
    nall[[3, 17]] = 0   # all-missing loci take the NumAlleles==0 branch (UpdateFst.c:46-48)
    ref.configure(N, L, K, A, FPRIORMEAN=0.05, FPRIORSD=0.07)
    m = oracle.make_model(N, L, K, A, fpriormean=0.05, fpriorsd=0.07)
    P = random_p(rng, L, K, A, nall)
    eps = np.zeros((L, A))
    for l in range(L):
        if nall[l]:
            eps[l, :nall[l]] = rng.dirichlet(np.ones(nall[l]))
    for f in rng.uniform(0.001, 0.99, 20):
        for pop in range(K):
            assert oracle.lib().orc_flike_freqs(oracle.C.byref(m), oracle.ip(nall), f,
                                                oracle.dp(eps), oracle.dp(P), pop) == \
                ref.lib().FlikeFreqs(f, ref.dp(eps), ref.dp(P), ref.ip(nall), pop)
        g = rng.uniform(0.001, 0.99)
        assert oracle.lib().orc_fprior_diff(oracle.C.byref(m), f, g) == ref.lib().FPriorDiff(f, g)


@pytest.mark.parametrize("onefst", [0, 1])
def test_update_fst_matches_reference(ref, onefst):
    rng = np.random.default_rng(61 + onefst)
    N, L, K, A = 12, 30, 3, 3
    _, nall = random_geno(rng, N, L, A)
    ref.configure(N, L, K, A, ONEFST=onefst, FPRIORMEAN=0.05, FPRIORSD=0.05)
    m = oracle.make_model(N, L, K, A, onefst=onefst, fpriormean=0.05, fpriorsd=0.05)
    M = oracle.C.byref(m)
    nacc = 0
    for trial in range(40):
        P = random_p(rng, L, K, A, nall, a=2.0)
        eps = np.zeros((L, A))
        for l in range(L):
            eps[l, :nall[l]] = rng.dirichlet(np.ones(nall[l]) * 2)
        F = np.full(K, rng.uniform(0.02, 0.2)) if onefst else rng.uniform(0.02, 0.2, K)
        queue = _rand_queue(rng, 400)
        n1 = 1 if onefst else K
        z, u, pos = np.zeros(K), np.zeros(K), 0
        for p in range(n1):
            ref.set_rand(queue[pos:])
            z[p] = ref.lib().snorm()
            pos += ref.lib().sref_rand_consumed()
            old = F[p]
            new = old + 0.05 * z[p]
            u[p] = queue[pos] / TWO31
            if 0.0 < new < 1.0:
                lpd = oracle.lib().orc_fprior_diff(M, new, old)
                for p2 in range(p, K if onefst else p + 1):
                    lpd += oracle.lib().orc_flike_freqs(M, oracle.ip(nall), new, oracle.dp(eps),
                                                        oracle.dp(P), p2)
                    lpd -= oracle.lib().orc_flike_freqs(M, oracle.ip(nall), old, oracle.dp(eps),
                                                        oracle.dp(P), p2)
                if lpd < 0.0:
                    pos += 1
        Fr = F.copy()
        ref.set_rand(queue)
        ref.lib().UpdateFst(ref.dp(eps), ref.dp(Fr), ref.dp(P), ref.ip(nall))
        assert ref.lib().sref_rand_consumed() == pos
        Fo = F.copy()
        oracle.lib().orc_update_fst_given(M, oracle.ip(nall), oracle.dp(eps), oracle.dp(P),
                                          oracle.dp(Fo), oracle.dp(z), oracle.dp(u))
        assert np.array_equal(Fo, Fr)
        nacc += int(not np.array_equal(Fo, F))
    assert 3 < nacc < 40


@pytest.mark.parametrize("freqscorr", [0, 1])
def test_init_freq_priors_and_sums_match_reference(ref, freqscorr):
    rng = np.random.default_rng(71)
    N, L, K, A = 15, 20, 3, 6
    geno, nall = random_geno(rng, N, L, A, miss=0.1)
    ref.configure(N, L, K, A, FREQSCORR=freqscorr, LAMBDA=0.8, FPRIORMEAN=0.03)
    m = oracle.make_model(N, L, K, A, freqscorr=freqscorr, lambda0=0.8, fpriormean=0.03)
    ch = oracle.Chain(m, geno, nall)
    ch.init()
    er = np.zeros((L, A))
    Fr = np.zeros(K)
    ref.lib().InitFreqPriors(ref.dp(er), ref.dp(Fr), ref.ip(geno_ref_layout(geno)),
                             ref.ip(nall))
    assert np.array_equal(ch.eps, er)
    assert np.array_equal(ch.F, Fr)
    # UpdateSums on the same state (output.c:654-707)
    ch.arrays["Q"][:] = random_q(rng, N, K)
    ch.arrays["P"][:] = random_p(rng, L, K, A, nall)
    ch.arrays["lambda_"][:] = rng.uniform(0.1, 2, K)
    ch.arrays["F"][:] = rng.uniform(0.01, 0.3, K)
    sums = {k: np.zeros_like(ch.arrays[k]) for k in ("QSum", "PSum", "FstSum", "epsSum",
                                                       "lambdaSum")}
    for _ in range(2):
        ref.lib().UpdateSums(ref.dp(ch.Q), ref.dp(sums["QSum"]), None, ref.dp(ch.P),
                             ref.dp(sums["PSum"]), ref.dp(ch.F), ref.dp(sums["FstSum"]),
                             ref.ip(nall), None, ref.dp(ch.eps), ref.dp(sums["epsSum"]),
                             ref.dp(ch.lambda_), ref.dp(sums["lambdaSum"]), None, None, 0)
        oracle.lib().orc_collect(oracle.C.byref(m), oracle.bp(ch.geno), oracle.ip(nall),
                                 oracle.C.byref(ch.state))
    for k, v in sums.items():
        assert np.array_equal(ch.arrays[k], v), k
    assert ch.state.nsum == 2
    like = ref.lib().CalcLike(ref.ip(geno_ref_layout(geno)), None, ref.dp(ch.Q), ref.dp(ch.P),
                              None, None, None)
    assert ch.like == like
    assert ch.sumlikes == 2 * like and ch.sumsqlikes == 2 * like * like


def test_initialize_z_rule_matches_reference(ref):
    # InitializeZ draws RandomInteger(0, K-1) = (int)(u*K) per non-missing copy in
    # (ind, line, loc) order (Init.c:142-163); the oracle applies the same rule to uz(w).
    rng = np.random.default_rng(81)
    N, L, K, A = 9, 13, 5, 3
    geno, nall = random_geno(rng, N, L, A, miss=0.2)
    ref.configure(N, L, K, A)
    queue = _rand_queue(rng, N * L * 2)
    Zr = np.zeros((N, 2, L), np.int32)
    ref.set_rand(queue)
    ref.lib().InitializeZ(ref.ip(geno_ref_layout(geno)), ref.individuals(N), ref.ip(Zr))
    expect = np.full((N, 2, L), -9, np.int32)
    pos = 0
    for i in range(N):
        for c in range(2):
            for l in range(L):
                if geno[i, c, l] != -9:
                    expect[i, c, l] = int(queue[pos] / TWO31 * K)
                    pos += 1
    assert np.array_equal(Zr, expect)
    ch = oracle.Chain(oracle.make_model(N, L, K, A, seed=4, chain=2), geno, nall)
    _, _, Zo = ch.init_z(want_z=True)
    for i in range(N):
        for l in range(L):
            w = oracle.philox_block(4, 2, 0, l, i >> 1, oracle.PH_INITZ << 24)
            j = 0
            for c in range(2):
                e = -9
                if geno[i, c, l] != -9:
                    e = int(w[2 * (i & 1) + j] * 2.0 ** -32 * K)
                    j += 1
                assert Zo[i, c, l] == e


def test_update_p_parameters_match_reference_in_distribution(ref):
    # The reference draws Dirichlet(par) with GKM gamma (ran.c:1094-1272), the
    # oracle with Marsaglia-Tsang: same distribution, so compare means over many
    # draws against the analytic mean par/sum(par) (UpdateP.c:158-174).
    rng = np.random.default_rng(91)
    N, L, K, A = 4, 3, 2, 4
    geno, nall = random_geno(rng, N, L, A, miss=0.0)
    nall[:] = [4, 3, 2]
    geno = np.minimum(geno, 1).astype(np.int8)
    for freqscorr in (1, 0):
        ref.configure(N, L, K, A, FREQSCORR=freqscorr)
        m = oracle.make_model(N, L, K, A, freqscorr=freqscorr)
        ch = oracle.Chain(m, geno, nall)
        Z = rng.integers(0, K, size=(N, 2, L)).astype(np.int32)
        nA = np.zeros((L, K, A), np.int32)
        ref.lib().GetNumFromPops(ref.ip(nA), ref.ip(geno_ref_layout(geno)), ref.ip(Z),
                                 ref.ip(nall), ref.individuals(N))
        eps = np.zeros((L, A))
        for l in range(L):
            eps[l, :nall[l]] = rng.dirichlet(np.ones(nall[l]))
        F = np.array([0.2, 0.05])
        lam = np.array([0.7, 1.3])
        par = np.zeros((L, K, A))
        for l in range(L):
            for k in range(K):
                for a in range(nall[l]):
                    par[l, k, a] = (eps[l, a] * (1 - F[k]) / F[k] if freqscorr else lam[k]) + \
                        nA[l, k, a]
        mean = par / np.maximum(par.sum(axis=2, keepdims=True), 1e-300)
        R = 3000
        Pr = np.zeros((L, K, A))
        Po = np.zeros((L, K, A))
        for r in range(R):
            ra = rng.random(L * A * K * 64)
            P = np.zeros((L, K, A))
            ref.lib().UpdateP(ref.dp(P), ref.dp(eps), ref.dp(F), ref.ip(nall),
                              ref.ip(geno_ref_layout(geno)), ref.ip(Z), ref.dp(lam),
                              ref.individuals(N), ref.dp(ra))
            Pr += P
            Po += ch.update_p(nA, eps, F, lam, r)
        Pr /= R
        Po /= R
        # Dirichlet marginal sd <= 0.5/sqrt(sum(par)+1); 5-sigma on the mean
        sd = 0.5 / np.sqrt(par.sum(axis=2, keepdims=True) + 1) / np.sqrt(R)
        assert np.all(np.abs(Pr - mean) < 5 * sd + 1e-12)
        assert np.all(np.abs(Po - mean) < 5 * sd + 1e-12)


def test_update_q_gibbs_parameters_match_reference_in_distribution(ref):
    # UpdateQAdmixture: Q[i] ~ Dirichlet(alpha + NumLociPop[i]) (UpdateQ.c:406-420)
    rng = np.random.default_rng(101)
    N, L, K, A = 3, 6, 3, 2
    geno, nall = random_geno(rng, N, L, A, miss=0.1, biallelic=True)
    ref.configure(N, L, K, A)
    Z = np.where(geno == -9, -9, rng.integers(0, K, size=geno.shape)).astype(np.int32)
    alpha = np.array([0.3, 1.0, 2.5])
    m = oracle.make_model(N, L, K, A)
    ch = oracle.Chain(m, geno, nall)
    nQ = np.zeros((N, K), np.int32)
    for i in range(N):
        row = np.zeros(K, np.int32)
        ref.lib().GetNumLociPop(ref.ip(row), ref.ip(Z), i)
        nQ[i] = row
    par = alpha + nQ
    mean = par / par.sum(axis=1, keepdims=True)
    R = 4000
    ref.set_rand(_rand_queue(rng, R * N * K * 40))
    Qr = np.zeros((N, K))
    Qo = np.zeros((N, K))
    for r in range(R):
        Q = np.zeros((N, K))
        ref.lib().UpdateQAdmixture(ref.dp(Q), ref.ip(Z), ref.dp(alpha), ref.individuals(N), None)
        Qr += Q
        Qo += ch.update_q_gibbs(nQ, alpha, r)
    assert ref.lib().sref_rand_overrun() == 0
    sd = 0.5 / np.sqrt(par.sum(axis=1, keepdims=True) + 1) / np.sqrt(R)
    assert np.all(np.abs(Qr / R - mean) < 5 * sd)
    assert np.all(np.abs(Qo / R - mean) < 5 * sd)


def test_est_log_prob_formula(ref):
    # "Estimated Ln Prob of Data" = mean - var/2 with SampleVar (output.c:457-467, mymath.c:52-62)
    s, ss, n = -12345.6, 152460000.0, 1000
    mean = s / n
    var = (ss - s * s / n) / (n - 1)
    assert ref.lib().EstLogProb(s, ss, n) == mean - var / 2.0


def _individuals_with_flags(ref, flags):
    """struct IND array (structure.h:52-59) with PopFlag set per individual."""
    N = len(flags)
    inds = ref.individuals(N)
    buf = bytearray(ref.IND_SIZE * N)
    for i, f in enumerate(flags):
        buf[i * ref.IND_SIZE + 16:i * ref.IND_SIZE + 20] = int(f).to_bytes(4, "little",
                                                                          signed=True)
    import ctypes
    ctypes.memmove(inds, bytes(buf), len(buf))
    return inds


@pytest.mark.parametrize("K,A", [(3, 5), (2, 2)])
def test_popflag_only_counts_match_reference(ref, K, A):
    """PFROMPOPFLAGONLY: GetNumFromPops (UpdateP.c:60-99) counts only individuals
    whose PopFlag is 1 (values 0, 1 and 2 present); the ancestry counts keep all."""
    rng = np.random.default_rng(300 + K)
    N, L = 21, 17
    geno, nall = random_geno(rng, N, L, A, miss=0.05)
    flags = rng.integers(0, 3, size=N).astype(np.int32)
    flags[:3] = [1, 0, 2]
    Q = random_q(rng, N, K)
    P = random_p(rng, L, K, A, nall)
    rep, seed, chain = 4, 99, 1
    ref.configure(N, L, K, A, PFROMPOPFLAGONLY=1)
    g32 = geno_ref_layout(geno)
    Zr = np.zeros((N, 2, L), np.int32)
    ref.lib().UpdateZ(ref.ip(Zr), ref.dp(Q), ref.dp(P), ref.ip(g32),
                      ref.dp(z_uniforms(seed, chain, rep, N, L)))
    nAr = np.zeros((L, K, A), np.int32)
    ref.lib().GetNumFromPops(ref.ip(nAr), ref.ip(g32), ref.ip(Zr), ref.ip(nall),
                             _individuals_with_flags(ref, flags))
    ref.configure(N, L, K, A)   # back to the default for the other tests
    m = oracle.make_model(N, L, K, A, seed=seed, chain=chain, pfrompopflagonly=1)
    ch = oracle.Chain(m, geno, nall, popflag=flags)
    nA, nQ, Z = ch.update_z(Q, P, rep, want_z=True)
    assert np.array_equal(Z, Zr)
    assert np.array_equal(nA, nAr)
    assert nA.sum() < (geno != -9).sum()                 # the filter removed copies
    assert nQ.sum() == (geno != -9).sum()                # ancestry counts keep them all
    nA2 = np.zeros_like(nA)
    nQ2 = np.zeros_like(nQ)
    oracle.lib().orc_count_from_z(oracle.C.byref(m), oracle.bp(ch.geno), oracle.ip(Zr),
                                  oracle.ip(nA2), oracle.ip(nQ2))
    assert np.array_equal(nA2, nAr)


def test_ancestdist_histogram_matches_reference(ref):
    """ANCESTDIST: UpdateSums (output.c:691-700) boxes every Q[i,k] among
    NUMBOXES boxes, Q = 1 into the last; the oracle's collection step gives the
    same histogram over several sweeps, including the edge values 0 and 1."""
    rng = np.random.default_rng(404)
    N, L, K, A, nb = 12, 9, 3, 4, 7
    geno, nall = random_geno(rng, N, L, A, miss=0.1)
    ref.configure(N, L, K, A, ANCESTDIST=1, NUMBOXES=nb)
    m = oracle.make_model(N, L, K, A, ancestdist=1, numboxes=nb, computeprob=0)
    ch = oracle.Chain(m, geno, nall)
    ch.init()
    ad = np.zeros((N, K, nb), np.int32)
    sums = {k: np.zeros_like(ch.arrays[k]) for k in ("QSum", "PSum", "FstSum", "epsSum",
                                                     "lambdaSum")}
    for sweep in range(5):
        Q = random_q(rng, N, K, a=0.4)
        if sweep == 0:
            Q[0] = [1.0, 0.0, 0.0]   # the Q = 1.000 and Q = 0 edges
        ch.arrays["Q"][:] = Q
        ref.lib().UpdateSums(ref.dp(ch.Q), ref.dp(sums["QSum"]), None, ref.dp(ch.P),
                             ref.dp(sums["PSum"]), ref.dp(ch.F), ref.dp(sums["FstSum"]),
                             ref.ip(nall), ref.ip(ad), ref.dp(ch.eps), ref.dp(sums["epsSum"]),
                             ref.dp(ch.lambda_), ref.dp(sums["lambdaSum"]), None, None, 0)
        oracle.lib().orc_collect(oracle.C.byref(m), oracle.bp(ch.geno), oracle.ip(nall),
                                 oracle.C.byref(ch.state))
    ref.configure(N, L, K, A)
    assert np.array_equal(ch.ancestDist, ad)
    assert ad.sum() == 5 * N * K and ad[0, 0, nb - 1] >= 1 and ad[0, 1, 0] >= 1
