"""Pin the oracle's own algorithms by mathematics: published known-answer
vectors, library routines, closed forms and invariants.  CPU only."""
import math

import numpy as np
import pytest
from scipy import stats

import oracle
from helpers import random_geno, random_p, random_q

# Random123 known-answer vectors for philox4x32-10 (counter, key, output)
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert list(oracle.philox(ctr, key)) == out


def test_uniform_conventions():
    assert oracle.lib().orc_uz(0) == 0.0
    assert oracle.lib().orc_uz(0xffffffff) < 1.0
    assert 0.0 < oracle.lib().orc_uo(0) and oracle.lib().orc_uo(0xffffffff) < 1.0


def test_lgamma_against_library():
    for z in np.concatenate([np.linspace(1e-3, 5, 200), 10 ** np.linspace(0.7, 7, 50)]):
        assert abs(oracle.lgamma(z) - math.lgamma(z)) <= 1e-12 * max(1.0, abs(math.lgamma(z)))


def test_normal_draws_are_standard_normal():
    x = np.array([oracle.normal(5, 1, r, i, 0, 6) for r in range(20) for i in range(250)])
    assert stats.kstest(x, "norm").pvalue > 1e-3
    assert abs(x.mean()) < 0.06 and abs(x.std() - 1) < 0.05


@pytest.mark.parametrize("a", [0.05, 0.3, 0.999, 1.0, 2.5, 40.0, 3000.0])
def test_gamma_draws_match_gamma_distribution(a):
    x = np.array([oracle.gamma(a, 11, 0, r, i, 3, 4) for r in range(8) for i in range(500)])
    assert stats.kstest(x, "gamma", args=(a,)).pvalue > 1e-3
    # and distinct indices give distinct draws (the counter addresses the stream)
    assert len(np.unique(x)) > 0.99 * len(x)


def test_z_draw_frequencies_follow_q_times_p():
    # closed form P(z = k) = Q_k P_{k,x} / sum_j Q_j P_{j,x}  (UpdateZ.c:44-50)
    N, L, K, A = 2, 1, 4, 2
    geno = np.zeros((N, 2, L), np.int8)
    geno[:, 1, 0] = 1
    nall = np.array([2], np.int32)
    Q = np.array([[0.1, 0.2, 0.3, 0.4], [0.7, 0.1, 0.1, 0.1]])
    P = np.array([[[0.2, 0.8], [0.5, 0.5], [0.9, 0.1], [0.4, 0.6]]])
    ch = oracle.Chain(oracle.make_model(N, L, K, A, seed=3), geno, nall)
    R = 6000
    counts = np.zeros((N, 2, K))
    for r in range(R):
        _, _, Z = ch.update_z(Q, P, r, want_z=True)
        for i in range(N):
            for c in range(2):
                counts[i, c, Z[i, c, 0]] += 1
    for i in range(N):
        for c in range(2):
            x = geno[i, c, 0]
            p = Q[i] * P[0, :, x]
            p /= p.sum()
            assert stats.chisquare(counts[i, c], R * p).pvalue > 1e-4


def test_z_special_cases():
    rng = np.random.default_rng(2)
    N, L, A = 9, 30, 3
    geno, nall = random_geno(rng, N, L, A, miss=0.2)
    P1 = random_p(rng, L, 1, A, nall)
    ch = oracle.Chain(oracle.make_model(N, L, 1, A), geno, nall)
    _, _, Z = ch.update_z(np.ones((N, 1)), P1, 4, want_z=True)
    assert np.all(Z[geno == -9] == -9) and np.all(Z[geno != -9] == 0)   # K = 1
    K = 4
    P = random_p(rng, L, K, A, nall)
    onehot = np.zeros((N, K))
    onehot[np.arange(N), np.arange(N) % K] = 1.0
    ch = oracle.Chain(oracle.make_model(N, L, K, A), geno, nall)
    _, nQ, Z = ch.update_z(onehot, P, 4, want_z=True)
    for i in range(N):   # one-hot Q: deterministic z
        assert np.all(Z[i][geno[i] != -9] == i % K)
    assert np.array_equal(nQ.sum(axis=1), (geno != -9).sum(axis=(1, 2)))


def test_count_invariants_and_like_brute_force():
    rng = np.random.default_rng(4)
    N, L, K, A = 25, 40, 3, 5
    geno, nall = random_geno(rng, N, L, A, miss=0.07)
    Q = random_q(rng, N, K)
    P = random_p(rng, L, K, A, nall)
    ch = oracle.Chain(oracle.make_model(N, L, K, A), geno, nall)
    nA, nQ, Z = ch.update_z(Q, P, 9, want_z=True)
    assert np.array_equal(nA.sum(axis=(1, 2)), (geno != -9).sum(axis=(0, 1)))
    assert np.array_equal(nQ.sum(axis=1), (geno != -9).sum(axis=(1, 2)))
    brute = 0.0
    for i in range(N):
        for c in range(2):
            for l in range(L):
                x = geno[i, c, l]
                if x != -9:
                    brute += math.log(float(np.dot(Q[i], P[l, :, x])))
    assert abs(ch.calc_like(Q, P) - brute) <= 1e-12 * abs(brute)


def test_dirichlet_updates_have_the_posterior_mean():
    rng = np.random.default_rng(6)
    N, L, K, A = 3, 2, 2, 3
    geno, nall = random_geno(rng, N, L, A, miss=0.0)
    nall[:] = 3
    ch = oracle.Chain(oracle.make_model(N, L, K, A, freqscorr=0), geno, nall)
    nA = rng.integers(0, 20, size=(L, K, A)).astype(np.int32)
    lam = np.array([0.5, 2.0])
    R = 3000
    acc = sum(ch.update_p(nA, np.zeros((L, A)), np.full(K, 0.5), lam, r) for r in range(R)) / R
    par = lam[None, :, None] + nA
    mean = par / par.sum(axis=2, keepdims=True)
    sd = 0.5 / np.sqrt(par.sum(axis=2, keepdims=True) + 1) / np.sqrt(R)
    assert np.all(np.abs(acc - mean) < 5 * sd)


def test_ancestdist_box_rule_by_brute_force():
    """The oracle's ANCESTDIST histogram is floor(Q * NUMBOXES), clamped to the
    last box, counted over the collected sweeps (checked without the reference)."""
    rng = np.random.default_rng(17)
    N, L, K, A, nb = 10, 6, 4, 3, 20
    geno, nall = random_geno(rng, N, L, A, miss=0.0)
    m = oracle.make_model(N, L, K, A, ancestdist=1, numboxes=nb, computeprob=0)
    ch = oracle.Chain(m, geno, nall)
    ch.init()
    expect = np.zeros((N, K, nb), np.int64)
    for s in range(6):
        Q = random_q(rng, N, K, a=0.3)
        Q[s % N] = np.eye(K)[s % K]
        ch.arrays["Q"][:] = Q
        oracle.lib().orc_collect(oracle.C.byref(m), oracle.bp(ch.geno), oracle.ip(ch.nall),
                                 oracle.C.byref(ch.state))
        box = np.minimum((Q * nb).astype(np.int64), nb - 1)
        for i in range(N):
            for k in range(K):
                expect[i, k, box[i, k]] += 1
    assert np.array_equal(ch.ancestDist, expect)


def test_popflag_filter_counts_by_recount():
    """PFROMPOPFLAGONLY in the oracle's UpdateZ: the allele counts equal a
    recount of the debug Z over the individuals with POPFLAG 1, the ancestry
    counts a recount over all (checked without the reference)."""
    rng = np.random.default_rng(23)
    N, L, K, A = 30, 25, 3, 4
    geno, nall = random_geno(rng, N, L, A, miss=0.05)
    flags = (rng.random(N) < 0.5).astype(np.int32)
    m = oracle.make_model(N, L, K, A, pfrompopflagonly=1, seed=3)
    ch = oracle.Chain(m, geno, nall, popflag=flags)
    nA, nQ, Z = ch.update_z(random_q(rng, N, K), random_p(rng, L, K, A, nall), 2, want_z=True)
    expect = np.zeros_like(nA)
    for i in range(N):
        if flags[i] != 1:
            continue
        for c in range(2):
            for l in range(L):
                if geno[i, c, l] != -9:
                    expect[l, Z[i, c, l], geno[i, c, l]] += 1
    assert np.array_equal(nA, expect)
    assert np.array_equal(nQ.sum(axis=1), (geno != -9).sum(axis=(1, 2)))
