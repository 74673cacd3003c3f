"""Replicate label alignment (structure_amd/host/align.c): the Hungarian
assignment against brute force over all permutations, and recovery of known
relabellings.  CPU only."""
import itertools

import numpy as np
import pytest

from structure_amd import host


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (3, 2), (5, 3), (6, 4), (7, 5)])
def test_assignment_is_optimal(n, seed):
    rng = np.random.default_rng(seed)
    for _ in range(10):
        c = rng.random((n, n)) * rng.choice([1.0, 100.0])
        if rng.random() < 0.3:
            c = np.round(c * 3)   # ties
        perm = host.assign(c)
        assert sorted(perm) == list(range(n))
        best = min(sum(c[i, p[i]] for i in range(n)) for p in itertools.permutations(range(n)))
        assert abs(sum(c[i, perm[i]] for i in range(n)) - best) <= 1e-9 * max(1.0, best)


def test_assignment_large_k_against_greedy_bound():
    rng = np.random.default_rng(9)
    n = 16
    c = rng.random((n, n))
    perm = host.assign(c)
    assert sorted(perm) == list(range(n))
    # a planted cheap diagonal under a permutation must be found
    sigma = rng.permutation(n)
    d = rng.random((n, n)) + 1.0
    d[np.arange(n), sigma] = 0.0
    assert np.array_equal(host.assign(d), sigma)


def test_replicates_relabelled_are_recovered():
    rng = np.random.default_rng(3)
    N, K, R = 60, 4, 5
    base = rng.dirichlet(np.full(K, 0.3), size=N)
    qbars, perms = [], []
    for r in range(R):
        sigma = rng.permutation(K) if r else np.arange(K)
        noisy = np.clip(base + rng.normal(0, 0.01, base.shape), 1e-6, None)
        noisy /= noisy.sum(axis=1, keepdims=True)
        q = np.empty_like(noisy)
        q[:, sigma] = noisy          # cluster a of the base is column sigma[a]
        qbars.append(q)
        perms.append(sigma)
    lnp = -1000.0 + rng.normal(0, 2, R)
    qmean, qsd, m, se, got = host.replicate_summary(np.array(qbars), lnp)
    for r in range(R):
        assert np.array_equal(got[r], perms[r])
    assert np.abs(qmean - base).max() < 0.02
    assert qsd.max() < 0.03
    assert abs(m - lnp.mean()) < 1e-9
    assert abs(se - lnp.std(ddof=1) / np.sqrt(R)) < 1e-9
