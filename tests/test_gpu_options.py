"""Model options of the sweep beyond the defaults, on the GPU against the oracle:

* PFROMPOPFLAGONLY (GetNumFromPops, UpdateP.c:85-88): the allele counts take
  only the individuals whose POPFLAG is 1, the ancestry counts all of them --
  in the fused UpdateZ of both genotype formats, its initial-Z variant, and
  whole chains (counts bit-exact, state within 1e-9).
* ANCESTDIST (UpdateSums, output.c:691-700): the per-(individual, pop)
  histogram of Q over the collected sweeps, fused into the Gibbs and
  Metropolis Q updates (bit-exact against the oracle's histogram).
"""
import numpy as np
import pytest

import oracle
import structure_amd as sa
from helpers import random_geno, random_p, random_q

pytestmark = pytest.mark.gpu


def _close(a, b, rtol):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    err = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    assert np.all((a == b) | (err <= rtol)), f"max rel err {err.max():.3e}"


def _data(fmt, N, L, seed):
    rng = np.random.default_rng(seed)
    if fmt == "snp":
        geno, nall = random_geno(rng, N, L, 2, miss=0.0, biallelic=True)
        geno = np.sort(geno, axis=1)   # heterozygotes as (0, 1)
        gone = rng.random((N, L)) < 0.05
        for c in range(2):
            geno[:, c, :] = np.where(gone, -9, geno[:, c, :])
        return geno, nall, 2, sa.GENO_SNP2
    geno, nall = random_geno(rng, N, L, 6, miss=0.05)
    return geno, nall, 6, sa.GENO_BYTE


@pytest.mark.parametrize("fmt,N,L,K", [("snp", 77, 600, 4), ("snp", 40, 9000, 3),
                                       ("byte", 45, 333, 5), ("byte", 130, 70, 3)])
def test_popflag_only_counts_bitexact(fmt, N, L, K):
    geno, nall, A, f = _data(fmt, N, L, seed=N + L)
    rng = np.random.default_rng(K)
    flags = rng.integers(0, 3, size=N).astype(np.int32)   # 0, 1 and 2: only 1 counts
    kw = dict(seed=21, chain=3, pfrompopflagonly=1)
    gch = sa.Chain.from_geno(sa.make_model(N, L, K, A, **kw), geno, nall, f, popflag=flags)
    och = oracle.Chain(oracle.make_model(N, L, K, A, **kw), geno, nall, popflag=flags)
    gch.init()
    och.init()
    assert np.array_equal(gch.get(sa.NA), och.nA)         # the initial-Z counts
    assert np.array_equal(gch.get(sa.NQ), och.nQ)
    Q, P = random_q(rng, N, K, 0.6), random_p(rng, L, K, A, nall)
    gch.set(sa.Q, Q)
    gch.set(sa.P, P)
    for rep in (3, 4):
        z = gch.update_z(rep, like=True, want_z=True)
        nA, nQ, Zo = och.update_z(Q, P, rep, want_z=True)
        assert np.array_equal(z.astype(np.int32), Zo)
        assert np.array_equal(gch.get(sa.NA), nA)
        assert np.array_equal(gch.get(sa.NQ), nQ)
        assert nA.sum() < (geno != -9).sum() and nQ.sum() == (geno != -9).sum()
        gch.collect(rep)
        _close(gch.get(sa.LIKE)[0], och.calc_like(Q, P), 1e-12)   # lnP keeps every copy
    gch.update_z(5)   # the plain variant (no lnP, no debug Z)
    assert np.array_equal(gch.get(sa.NA), och.update_z(Q, P, 5)[0])


def test_popflag_without_data_flags_is_refused():
    geno, nall, A, f = _data("snp", 20, 40, seed=1)
    with pytest.raises(sa.SgpuError) as e:
        sa.Chain.from_geno(sa.make_model(20, 40, 2, A, pfrompopflagonly=1), geno, nall, f)
    assert e.value.code == sa.ERR_ARG


@pytest.mark.parametrize("fmt", ["snp", "byte"])
def test_options_whole_chain_follows_oracle(fmt):
    """PFROMPOPFLAGONLY and ANCESTDIST together over Metropolis and Gibbs sweeps
    replayed from graphs: counts, histogram and state follow the oracle."""
    N, L, K, nb = 60, 300, 3, 40
    geno, nall, A, f = _data(fmt, N, L, seed=7)
    flags = (np.arange(N) % 3 != 0).astype(np.int32)
    kw = dict(freqscorr=1, inferalpha=1, inferlambda=1, burnin=3, metrofreq=2, seed=5,
              chain=2, pfrompopflagonly=1, ancestdist=1, numboxes=nb)
    gch = sa.Chain.from_geno(sa.make_model(N, L, K, A, **kw), geno, nall, f, popflag=flags)
    och = oracle.Chain(oracle.make_model(N, L, K, A, **kw), geno, nall, popflag=flags)
    gch.init()
    och.init()
    gch.run(0, 11)
    for rep in range(11):
        och.sweep(rep)
    assert np.array_equal(gch.get(sa.NA), och.nA)
    assert np.array_equal(gch.get(sa.NQ), och.nQ)
    ad = gch.get(sa.ANCESTDIST)
    assert ad.sum() == 8 * N * K
    assert np.array_equal(ad, och.ancestDist)
    for sel, ref in [(sa.Q, och.Q), (sa.P, och.P), (sa.QSUM, och.QSum), (sa.PSUM, och.PSum),
                     (sa.ALPHA, och.alpha), (sa.EPS, och.eps)]:
        _close(gch.get(sel), ref, 1e-9)


def test_ancestdist_in_a_batch_and_state_access():
    """Chains with and without ANCESTDIST in one batch: the histogram exists
    only for the chains that asked for it, and reads back what was written."""
    N, L = 50, 200
    geno, nall, A, f = _data("snp", N, L, seed=3)
    fmt, words = sa.pack(geno, f)
    models = [sa.make_model(N, L, 2, A, ancestdist=1, numboxes=10, burnin=0, seed=1, chain=0),
              sa.make_model(N, L, 4, A, burnin=0, seed=1, chain=1)]
    b = sa.Batch(models, [(fmt, words, nall)])
    b.init()
    b.run(0, 6)
    ad = b.get_chain(0, sa.ANCESTDIST)
    assert ad.shape == (N, 2, 10) and ad.sum() == 6 * N * 2
    assert np.all(ad.sum(axis=2) == 6)
    with pytest.raises(sa.SgpuError):
        b.get_chain(1, sa.ANCESTDIST)
    b.set_chain(0, sa.ANCESTDIST, np.zeros_like(ad))
    assert b.get_chain(0, sa.ANCESTDIST).sum() == 0
    b.close()
