"""Chain-level agreement with the reference's own CPU sweep: over 8 seeds,
posterior-mean Q (after label alignment), estimated ln P(D) and the mean
alpha / F / lambda of the oracle chains (here) and of the GPU chains (batched,
on the GPU) agree with the reference chains within Monte Carlo error.  The
reference statistics are fixtures (tests/golden/chains_*.npz) written by
tests/golden/make_golden_chains.py from oracle/_ref."""
import os

import numpy as np
import pytest

import oracle
from chain_cases import CASES, SEEDS, compare

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture(case):
    return np.load(os.path.join(GOLD, f"chains_{case}.npz"))


def _model_kw(c, seed):
    o = dict(c["opts"])
    o["alpha0"] = o.pop("alpha")
    o["lambda0"] = o.pop("lambda_")
    return dict(o, burnin=c["burnin"], numreps=c["numreps"], seed=seed, chain=c["K"])


def _stats_from(sums, c):
    n = c["numreps"]
    q, lk, al, fs, la = sums
    mean = lk[1] / n
    var = (lk[2] - lk[1] ** 2 / n) / (n - 1)
    return dict(qbar=q / n, lnp=mean - var / 2, alpha=al / n, fst=fs / n, lam=la / n)


def _check(ref, got, c):
    r = compare(ref, got, c["K"])
    assert r["lnp_z"] < 4.0, r
    assert r["alpha_z"] < 4.0, r
    if c["opts"]["freqscorr"]:
        assert r["fst_z"] < 4.0, r
    if c["opts"]["inferlambda"]:
        assert r["lam_z"] < 4.0, r
    assert r["q_frac_outside"] < 0.02, r
    return r


@pytest.mark.parametrize("case", list(CASES))
def test_oracle_chains_match_reference_chains(case):
    c = CASES[case]
    g = _fixture(case)
    geno, nall, A = g["geno"], g["nall"], int(g["A"])
    stats = dict(qbar=[], lnp=[], alpha=[], fst=[], lam=[])
    for s in SEEDS:
        ch = oracle.Chain(oracle.make_model(c["N"], c["L"], c["K"], A, **_model_kw(c, s)),
                          geno, nall)
        ch.init()
        for rep in range(c["burnin"] + c["numreps"]):
            ch.sweep(rep)
        st = _stats_from((ch.QSum, [ch.like, ch.sumlikes, ch.sumsqlikes], ch.alphaSum,
                          ch.FstSum, ch.lambdaSum), c)
        for k in stats:
            stats[k].append(st[k])
    _check({k: g[k] for k in stats}, stats, c)


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
def test_gpu_chains_match_reference_chains(case):
    import structure_amd as sa
    c = CASES[case]
    g = _fixture(case)
    geno, nall, A = g["geno"], g["nall"], int(g["A"])
    fmt, words = sa.pack(geno)
    models = [sa.make_model(c["N"], c["L"], c["K"], A, **_model_kw(c, s)) for s in SEEDS]
    b = sa.Batch(models, [(fmt, words, nall)])
    b.init()
    b.run(0, c["burnin"] + c["numreps"])
    stats = dict(qbar=[], lnp=[], alpha=[], fst=[], lam=[])
    for i in range(len(SEEDS)):
        st = _stats_from((b.get_chain(i, sa.QSUM), b.get_chain(i, sa.LIKE),
                          b.get_chain(i, sa.ALPHASUM), b.get_chain(i, sa.FSTSUM),
                          b.get_chain(i, sa.LAMBDASUM)), c)
        for k in stats:
            stats[k].append(st[k])
    _check({k: g[k] for k in stats}, stats, c)
