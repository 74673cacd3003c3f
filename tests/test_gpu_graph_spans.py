"""Sweeps replayed from graphs that span several sweeps (sgpu_run_sweeps):
whole Metropolis periods from a Metropolis sweep (periods up to 64 sweeps),
16-sweep spans without one when the period is longer or there is none, and
single-sweep graphs around them, across the burn-in boundary and over several run calls.  The replays
launch the same kernels as direct sweeps (sgpu_sweep), so every state array,
sum and lnP(D) is bit-identical."""
import numpy as np
import pytest

import structure_amd as sa
from helpers import random_geno

pytestmark = pytest.mark.gpu

STATE = (sa.Q, sa.P, sa.EPS, sa.FST, sa.ALPHA, sa.LAMBDA, sa.NA, sa.NQ, sa.QSUM, sa.PSUM,
         sa.EPSSUM, sa.FSTSUM, sa.ALPHASUM, sa.LAMBDASUM, sa.NSUM, sa.LIKE)

# (metrofreq, burnin, run calls as (rep0, n)): period graphs of 18 sweeps from
# Metropolis sweeps (6, 24, ...); one 20-sweep period per graph; a 70-sweep
# period, so 16-sweep Gibbs spans between Metropolis sweeps; no Metropolis
# sweeps, burn-in spans and collecting spans
CASES = {
    "period3": (3, 5, [(0, 7), (7, 40), (47, 13)]),
    "period20": (20, 10, [(0, 30), (30, 45)]),
    "period70": (70, 10, [(0, 30), (30, 60), (90, 5)]),
    "gibbs_only": (0, 17, [(0, 50)]),
}


def _chain(mf, burnin, fmt, seed):
    rng = np.random.default_rng(seed)
    if fmt == "snp":
        geno, nall = random_geno(rng, 48, 400, 2, miss=0.0, biallelic=True)
        geno = np.sort(geno, axis=1)   # heterozygotes as (0, 1)
        gone = rng.random((48, 400)) < 0.03   # missing genotypes: both copies
        geno[:, 0, :] = np.where(gone, -9, geno[:, 0, :])
        geno[:, 1, :] = np.where(gone, -9, geno[:, 1, :])
        A, f = 2, sa.GENO_SNP2
    else:
        geno, nall = random_geno(rng, 40, 150, 6, miss=0.03)
        A, f = 6, sa.GENO_BYTE
    m = sa.make_model(geno.shape[0], geno.shape[2], 3, A, burnin=burnin, metrofreq=mf,
                      inferalpha=1, inferlambda=1, freqscorr=1, seed=13, chain=2)
    return sa.Chain.from_geno(m, geno, nall, f)


@pytest.mark.parametrize("fmt", ["snp", "byte"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_span_graphs_equal_direct_sweeps(case, fmt):
    mf, burnin, calls = CASES[case]
    print(f"graph spans {case} {fmt}", flush=True)
    g = _chain(mf, burnin, fmt, seed=3)
    d = _chain(mf, burnin, fmt, seed=3)
    g.init()
    d.init()
    for rep0, n in calls:
        g.run(rep0, n)
        for rep in range(rep0, rep0 + n):
            d.sweep(rep)
        for sel in STATE:
            assert np.array_equal(g.get(sel), d.get(sel)), (case, fmt, rep0, sel)
    g.close()
    d.close()
