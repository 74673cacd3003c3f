"""Whole sweeps under the model options that decide which launch of a sweep
carries a phase, on the GPU against the oracle (structure.c:497-638 order):

* lambda inferred without the F model (log sums of the new P beside UpdateQ,
  decision in the epsilon launch), pooled and per pop (UpdateLambda.c:8-102);
* lambda inferred with the F model, per pop (its own launches after UpdateZ);
* ONEFST (one Fst proposal, UpdateFst.c:142-204) with lambda and the F model;
* a chain with more than 32 locus tiles, whose Gibbs update reads the ancestry
  counts folded in the previous sweep's epsilon launch;
* microsatellites with lambda inferred.

Each case runs direct launches (sgpu_sweep) and graph replays
(sgpu_run_sweeps) over burn-in and collected sweeps with Metropolis and Gibbs
Q updates.  Bars as in the other parity tests: counts bit-exact, state and
sums within 1e-9 relative.
"""
import numpy as np
import pytest

import oracle
import structure_amd as sa
from helpers import random_geno

pytestmark = pytest.mark.gpu

SWEEPS = 12


def _close(a, b, rtol=1e-9):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape
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
    geno, nall = random_geno(rng, N, L, 7, miss=0.05)
    return geno, nall, 7, sa.GENO_BYTE


CASES = {
    "lambda_no_f": ("snp", 64, 300, 3, dict(freqscorr=0, inferlambda=1)),
    "poplambda_no_f": ("snp", 64, 300, 3, dict(freqscorr=0, inferlambda=1, popspecificlambda=1)),
    "poplambda_f": ("snp", 64, 300, 4, dict(freqscorr=1, inferlambda=1, popspecificlambda=1)),
    "onefst_lambda": ("snp", 64, 300, 4, dict(freqscorr=1, onefst=1, inferlambda=1)),
    "fold_f_lambda": ("snp", 40, 9000, 3, dict(freqscorr=1, inferlambda=1)),
    "microsat_lambda": ("byte", 50, 120, 3, dict(freqscorr=0, inferlambda=1)),
}


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("case", sorted(CASES))
def test_sweep_options_follow_the_oracle(case, graphs):
    fmt, N, L, K, opts = CASES[case]
    geno, nall, A, f = _data(fmt, N, L, seed=11 + sorted(CASES).index(case))
    kw = dict(burnin=4, metrofreq=3, inferalpha=1, seed=77, chain=3, **opts)
    gch = sa.Chain.from_geno(sa.make_model(N, L, K, A, **kw), geno, nall, f)
    och = oracle.Chain(oracle.make_model(N, L, K, A, **kw), geno, nall)
    print(f"case {case} graphs={graphs}: N={N} L={L} K={K} fmt={fmt}", flush=True)
    gch.init()
    och.init()
    if graphs:
        gch.run(0, SWEEPS)
    else:
        for rep in range(SWEEPS):
            gch.sweep(rep)
    for rep in range(SWEEPS):
        och.sweep(rep)
    assert np.array_equal(gch.get(sa.NA), och.nA)
    assert np.array_equal(gch.get(sa.NQ), och.nQ)
    for sel, ref in [(sa.Q, och.Q), (sa.P, och.P), (sa.ALPHA, och.alpha),
                     (sa.LAMBDA, och.lambda_), (sa.QSUM, och.QSum), (sa.PSUM, och.PSum),
                     (sa.ALPHASUM, och.alphaSum), (sa.LAMBDASUM, och.lambdaSum)]:
        _close(gch.get(sel), ref)
    if opts.get("freqscorr", 1):
        for sel, ref in [(sa.EPS, och.eps), (sa.FST, och.F), (sa.EPSSUM, och.epsSum),
                         (sa.FSTSUM, och.FstSum)]:
            _close(gch.get(sel), ref)
    _close(gch.get(sa.LIKE), [och.like, och.sumlikes, och.sumsqlikes])
    assert gch.get(sa.NSUM)[0] == och.state.nsum == SWEEPS - 4
    # lambda moved at least once, so the decision ran on real sums
    assert not np.all(gch.get(sa.LAMBDA) == 1.0)
    gch.close()


@pytest.mark.parametrize("graphs", [False, True])
def test_host_write_of_eps_between_sweeps(graphs):
    """F-model lambda decides from the log-eps sums that the last epsilon pass
    left; after a host write of eps the next sweep must sum the written eps.
    Runs Gibbs and Metropolis sweeps on both sides of the write, with a count
    fold (more than 32 locus tiles) on the chain, against the oracle."""
    N, L, K = 40, 9000, 3
    geno, nall, A, f = _data("snp", N, L, seed=5)
    kw = dict(burnin=3, metrofreq=4, inferalpha=1, seed=91, chain=1, freqscorr=1, inferlambda=1)
    gch = sa.Chain.from_geno(sa.make_model(N, L, K, A, **kw), geno, nall, f)
    och = oracle.Chain(oracle.make_model(N, L, K, A, **kw), geno, nall)
    print(f"host eps write, graphs={graphs}", flush=True)
    gch.init()
    och.init()
    n1, n2 = 6, 7

    def run(r0, n):
        if graphs:
            gch.run(r0, n)
        else:
            for rep in range(r0, r0 + n):
                gch.sweep(rep)
        for rep in range(r0, r0 + n):
            och.sweep(rep)

    run(0, n1)
    rng = np.random.default_rng(8)
    eps = np.zeros((L, A))
    for l in range(L):
        if nall[l] > 0:
            eps[l, :nall[l]] = rng.dirichlet(np.ones(nall[l]))
    gch.set(sa.EPS, eps)
    och.eps[:] = eps
    run(n1, n2)
    assert np.array_equal(gch.get(sa.NA), och.nA)
    assert np.array_equal(gch.get(sa.NQ), och.nQ)
    for sel, ref in [(sa.Q, och.Q), (sa.P, och.P), (sa.EPS, och.eps), (sa.FST, och.F),
                     (sa.LAMBDA, och.lambda_), (sa.LAMBDASUM, och.lambdaSum)]:
        _close(gch.get(sel), ref)
    assert not np.all(gch.get(sa.LAMBDA) == 1.0)
    gch.close()
