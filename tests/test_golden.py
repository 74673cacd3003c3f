"""The oracle against fixtures produced by the reference's own C code
(tests/golden/ref_vectors.npz, written by tests/golden/make_golden.py from
oracle/_ref).  Runs without the reference sources."""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.npz")
TWO31 = 2.0 ** 31


@pytest.fixture(scope="module")
def g():
    return np.load(GOLD)   # plain arrays only (allow_pickle stays False)


def test_update_z_counts_like(g):
    geno, nall = g["z_geno"], g["z_nall"]
    N, _, L = geno.shape
    K, A = g["z_Q"].shape[1], g["z_P"].shape[2]
    seed, chain, rep = (int(x) for x in g["z_key"])
    m = oracle.make_model(N, L, K, A, seed=seed, chain=chain)
    ch = oracle.Chain(m, geno, nall)
    nA, nQ, Z = ch.update_z(g["z_Q"], g["z_P"], rep, want_z=True)
    assert np.array_equal(Z, g["z_Z"])
    assert np.array_equal(nA, g["z_nA"])
    assert np.array_equal(nQ, g["z_nQ"])
    assert ch.calc_like(g["z_Q"], g["z_P"]) == float(g["z_like"])
    for i in range(N):
        d = oracle.lib().orc_calc_like_ind_diff(
            oracle.C.byref(m), oracle.bp(ch.geno), oracle.dp(np.ascontiguousarray(g["z_Q2"][i])),
            oracle.dp(np.ascontiguousarray(g["z_Q"][i])), oracle.dp(g["z_P"]), i)
        assert d == g["z_diffs"][i]


def test_lgamma(g):
    for z, v in zip(g["lg_z"], g["lg_v"]):
        assert oracle.lgamma(z) == v


def test_alpha_helpers(g):
    Q = g["lp_Q"]
    N, K = Q.shape
    m = oracle.make_model(N, 3, K, 2, alphapriora=1.7, alphapriorb=2.5)
    M = oracle.C.byref(m)
    for j, a in enumerate(g["lp_alphas"]):
        assert oracle.lib().orc_log_prob_q(M, oracle.dp(Q), a) == g["lp_logprobq"][j]
        for p in range(K):
            assert oracle.lib().orc_log_prob_q_onepop(M, oracle.dp(Q), a, 3.3, p) == \
                g["lp_onepop"][j, p]
        assert oracle.lib().orc_alpha_prior_diff(M, a, 0.9) == g["lp_prior"][j]


def test_fst_helpers(g):
    P, eps, nall = g["fl_P"], g["fl_eps"], g["fl_nall"]
    L, K, A = P.shape
    m = oracle.make_model(10, L, K, A, fpriormean=0.05, fpriorsd=0.07)
    M = oracle.C.byref(m)
    for j, f in enumerate(g["fl_f"]):
        for p in range(K):
            assert oracle.lib().orc_flike_freqs(M, oracle.ip(nall), f, oracle.dp(eps),
                                                oracle.dp(P), p) == g["fl_v"][j, p]
        assert oracle.lib().orc_fprior_diff(M, f, 0.04) == g["fl_prior"][j]


def test_update_alpha_outcomes(g):
    Q = g["lp_Q"]
    N, K = Q.shape
    m = oracle.make_model(N, 3, K, 2, alphapropsd=0.4, alphamax=3.0)
    moved = 0
    for a0, a1, z, u in zip(g["al_in"], g["al_out"], g["al_z"], g["al_u"]):
        a = a0.copy()
        zz = np.full(K, z)
        uu = np.full(K, u)
        oracle.lib().orc_update_alpha_given(oracle.C.byref(m), oracle.dp(Q), oracle.dp(a),
                                            oracle.dp(zz), oracle.dp(uu))
        assert np.array_equal(a, a1)
        moved += int(not np.array_equal(a1, a0))
    assert 0 < moved < len(g["al_in"])


def test_update_eps_outcome(g):
    P, eps, nall, F, q = g["fl_P"], g["fl_eps"], g["fl_nall"], g["ep_F"], g["ep_queue"]
    L, K, A = P.shape
    N = 10
    invs = N ** -0.5
    a1 = np.zeros(L, np.int32)
    a2 = np.zeros(L, np.int32)
    d = np.zeros(L)
    u = np.zeros(L)
    pos = 0
    for l in range(L):
        na = nall[l]
        if na > 1:   # RandomInteger x2, RandomReal, then rnd() if in range (UpdateEpsilon.c:84-138)
            a1[l] = int(q[pos] / TWO31 * na)
            a2[l] = int(q[pos + 1] / TWO31 * (na - 1))
            a2[l] += a2[l] >= a1[l]
            d[l] = 0.0 + (q[pos + 2] / TWO31) * (invs - 0.0)
            pos += 3
            if eps[l, a1[l]] + d[l] < 1.0 and eps[l, a2[l]] - d[l] > 0.0:
                u[l] = q[pos] / TWO31
                pos += 1
    assert pos == int(g["ep_consumed"])
    m = oracle.make_model(N, L, K, A)
    e = eps.copy()
    oracle.lib().orc_update_eps_given(oracle.C.byref(m), oracle.ip(nall), oracle.dp(P),
                                      oracle.dp(F), 0.8, oracle.dp(e), oracle.ip(a1),
                                      oracle.ip(a2), oracle.dp(d), oracle.dp(u))
    assert np.array_equal(e, g["ep_out"])


def test_init_freq_priors(g):
    geno, nall = g["z_geno"], g["z_nall"]
    N, _, L = geno.shape
    K, A = g["z_Q"].shape[1], g["z_P"].shape[2]
    m = oracle.make_model(N, L, K, A, freqscorr=1, lambda0=0.9, fpriormean=0.02)
    ch = oracle.Chain(m, geno, nall)
    ch.init()
    assert np.array_equal(ch.eps, g["ifp_eps"])
    assert np.array_equal(ch.F, g["ifp_F"])
