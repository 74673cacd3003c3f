"""Gibbs Q update through the two-level ancestry-count reduction: chains with
more than 32 UpdateZ locus tiles (L > 8192 loci of 2-bit words) have their
per-tile partials folded by k_fold_rows before the Gibbs draw, chains with
fewer are summed by the Gibbs threads directly.  Both paths, alone and mixed in
one batch, against the oracle's UpdateQAdmixture on the counts of the same
UpdateZ (counts bit-exact, Q within 1e-12 relative)."""
import numpy as np
import pytest

import oracle
import structure_amd as sa
from helpers import random_p, random_q

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _close(a, b):
    err = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    assert np.all((a == b) | (err <= RTOL)), f"max rel err {err.max():.3e}"


def _data(N, L, seed):
    d = sa.synth_snp(N, L, Ktrue=3, qalpha=0.5, miss=0.02, seed=seed, want_geno=True)
    return d


@pytest.mark.parametrize("L", [9000, 20000])
def test_gibbs_after_fold_matches_oracle(L):
    N, K = 40, 4
    print(f"fold path N={N} L={L} K={K}", flush=True)
    rng = np.random.default_rng(L)
    d = _data(N, L, seed=L)
    m = dict(seed=13, chain=2, metrofreq=10)
    gch = sa.Chain(sa.make_model(N, L, K, 2, **m), sa.GENO_SNP2, d["words"], d["nall"])
    och = oracle.Chain(oracle.make_model(N, L, K, 2, **m), d["geno"], d["nall"])
    gch.init()
    Q = random_q(rng, N, K, 0.7)
    P = random_p(rng, L, K, 2, d["nall"])
    alpha = np.array([0.5, 1.1, 0.9, 2.0])
    gch.set(sa.Q, Q)
    gch.set(sa.P, P)
    gch.set(sa.ALPHA, alpha)
    gch.update_z(5)
    _, nQ = och.update_z(Q, P, 5)
    assert np.array_equal(gch.get(sa.NQ), nQ)
    gch.update_q(6)   # rep 6 is not a Metropolis sweep at METROFREQ=10
    _close(gch.get(sa.Q), och.update_q_gibbs(nQ, alpha, 6))
    gch.close()


def test_batch_mixing_fold_and_direct_chains():
    """One chain above 32 locus tiles (fold) and two below (direct sums) in one
    batch: the fold launch covers only the first; every chain's Q matches."""
    K = 3
    shapes = [(36, 9500), (50, 700), (20, 2000)]
    datas = [_data(N, L, seed=100 + j) for j, (N, L) in enumerate(shapes)]
    models = [sa.make_model(N, L, K, 2, seed=7, chain=j, metrofreq=10)
              for j, (N, L) in enumerate(shapes)]
    b = sa.Batch(models, [(sa.GENO_SNP2, d["words"], d["nall"]) for d in datas],
                 list(range(len(shapes))))
    b.init()
    rng = np.random.default_rng(3)
    alpha = np.array([0.6, 1.4, 1.0])
    states = []
    for j, ((N, L), d) in enumerate(zip(shapes, datas)):
        Q = random_q(rng, N, K)
        P = random_p(rng, L, K, 2, d["nall"])
        b.set_chain(j, sa.Q, Q)
        b.set_chain(j, sa.P, P)
        b.set_chain(j, sa.ALPHA, alpha)
        states.append((Q, P))
    b.update_z(11)
    b.update_q(12)
    for j, ((N, L), d) in enumerate(zip(shapes, datas)):
        och = oracle.Chain(oracle.make_model(N, L, K, 2, seed=7, chain=j), d["geno"], d["nall"])
        Q, P = states[j]
        _, nQ = och.update_z(Q, P, 11)
        assert np.array_equal(b.get_chain(j, sa.NQ), nQ)
        _close(b.get_chain(j, sa.Q), och.update_q_gibbs(nQ, alpha, 12))
    b.close()
