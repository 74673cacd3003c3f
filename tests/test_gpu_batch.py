"""Chain batches (sgpu_create_batch): chains of uneven K, flags, seeds and data
sets advanced by one launch per kernel must follow exactly the path each chain
follows alone.  Each chain's Philox stream depends only on its own key, and the
counts are exact integers whatever the tiling, so Z, the counts, P, Q, alpha,
lambda, eps, F and the posterior sums are bit-identical; lnP(D) is an FP64 sum
over the UpdateZ tiles, whose number depends on the batch, so it agrees to
1e-12 relative."""
import numpy as np
import pytest

import structure_amd as sa

pytestmark = pytest.mark.gpu

EXACT = (sa.Q, sa.P, sa.EPS, sa.FST, sa.ALPHA, sa.LAMBDA, sa.NA, sa.NQ, sa.QSUM, sa.PSUM,
         sa.EPSSUM, sa.FSTSUM, sa.ALPHASUM, sa.LAMBDASUM, sa.NSUM)


def _datasets():
    snp = sa.synth_snp(90, 700, Ktrue=4, qalpha=0.5, miss=0.02, seed=21)
    ms = sa.synth_microsat(60, 300, amin=2, amax=7, Ktrue=3, miss=0.03, seed=22)
    fmt, words = sa.pack(ms["geno"], sa.GENO_BYTE)
    return [(sa.GENO_SNP2, snp["words"], snp["nall"], 90, 700, 2),
            (fmt, words, ms["nall"], 60, 300, ms["A"])]


CHAINS = [  # (data set, K, options)
    (0, 1, dict(seed=1, chain=0)),
    (0, 2, dict(seed=1, chain=1, inferlambda=1)),
    (0, 5, dict(seed=7, chain=0, popalphas=1)),
    (0, 3, dict(seed=3, chain=4, freqscorr=0, inferlambda=1, popspecificlambda=1)),
    (0, 5, dict(seed=8, chain=2, onefst=1)),
    (1, 3, dict(seed=5, chain=0, freqscorr=0, inferalpha=1)),
    (1, 4, dict(seed=6, chain=1, inferlambda=1)),
]


def _models(dsets, common, chains=CHAINS):
    out = []
    for di, K, kw in chains:
        _, _, _, N, L, A = dsets[di]
        o = dict(common)
        o.update(kw)
        out.append(sa.make_model(N, L, K, A, **o))
    return out


def _batch_vs_alone(dsets, chains, graphs, nsweeps=17):
    common = dict(burnin=6, metrofreq=4)
    models = _models(dsets, common, chains)
    batch = sa.Batch(models, [d[:3] for d in dsets], [c[0] for c in chains])
    assert len(batch) == len(chains)
    batch.init()
    if graphs:
        batch.run(0, 9)
        batch.run(9, nsweeps - 9)
    else:
        for r in range(nsweeps):
            batch.sweep(r)
    for ci, (di, K, kw) in enumerate(chains):
        fmt, words, nall = dsets[di][:3]
        alone = sa.Chain(_models(dsets, common, chains)[ci], fmt, words, nall)
        alone.init()
        alone.run(0, nsweeps)
        for sel in EXACT:
            assert np.array_equal(batch.get_chain(ci, sel), alone.get(sel)), (ci, sel)
        lb, la = batch.get_chain(ci, sa.LIKE), alone.get(sa.LIKE)
        np.testing.assert_allclose(lb, la, rtol=1e-12)
        assert la[1] != 0.0
        alone.close()
    batch.close()


@pytest.mark.parametrize("graphs", [True, False])
def test_batch_equals_chains_run_alone(graphs):
    _batch_vs_alone(_datasets(), CHAINS, graphs)


# Chains whose sweeps place phases in different launches: lambda per pop under
# the F model (its own launches after UpdateZ), lambda without the F model
# (decided in the epsilon launch), more than 32 UpdateZ locus tiles (the Gibbs
# update reads the count fold of the previous epsilon launch), no F model and
# no lambda (epsilon blocks absent), one Fst with pop alphas.
MIXED = [
    (2, 3, dict(seed=2, chain=0, inferlambda=1, popspecificlambda=1)),
    (0, 4, dict(seed=2, chain=1, inferlambda=1, popspecificlambda=1)),
    (0, 2, dict(seed=4, chain=2, freqscorr=0)),
    (2, 2, dict(seed=5, chain=3, freqscorr=0, inferlambda=1)),
    (1, 3, dict(seed=6, chain=4, onefst=1, inferalpha=1, popalphas=1)),
    (2, 4, dict(seed=9, chain=5, inferalpha=1)),
]


@pytest.mark.parametrize("graphs", [True, False])
def test_batch_of_mixed_launch_roles_equals_chains_run_alone(graphs):
    big = sa.synth_snp(40, 9000, Ktrue=3, qalpha=0.5, miss=0.02, seed=23)
    dsets = _datasets() + [(sa.GENO_SNP2, big["words"], big["nall"], 40, 9000, 2)]
    print(f"mixed batch graphs={graphs}: {len(MIXED)} chains", flush=True)
    _batch_vs_alone(dsets, MIXED, graphs, nsweeps=14)


def test_batch_tuning_changes_tiles_not_results():
    dsets = _datasets()
    common = dict(burnin=3, metrofreq=3)
    models = _models(dsets, common)
    a = sa.Batch(models, [d[:3] for d in dsets], [c[0] for c in CHAINS])
    b = sa.Batch(models, [d[:3] for d in dsets], [c[0] for c in CHAINS],
                 tuning=sa.Tuning(z_blocks=1, metro_blocks=1, byte_loci=32, byte_chunk=3))
    c = sa.Batch(models, [d[:3] for d in dsets], [c[0] for c in CHAINS],
                 tuning=sa.Tuning(z_blocks=100000, metro_blocks=100000, byte_loci=256,
                                  byte_chunk=1))
    for x in (a, b, c):
        x.init()
        x.run(0, 8)
    for ci in range(len(CHAINS)):
        for other in (b, c):
            for sel in EXACT:
                assert np.array_equal(a.get_chain(ci, sel), other.get_chain(ci, sel)), (ci, sel)
            np.testing.assert_allclose(a.get_chain(ci, sa.LIKE), other.get_chain(ci, sa.LIKE),
                                       rtol=1e-12)


def test_batch_state_is_per_chain():
    dsets = _datasets()
    models = _models(dsets, dict(burnin=0, metrofreq=0))
    batch = sa.Batch(models, [d[:3] for d in dsets], [c[0] for c in CHAINS])
    batch.init()
    rng = np.random.default_rng(3)
    K2 = models[2].K
    q = rng.dirichlet(np.ones(K2), size=models[2].N)
    batch.set_chain(2, sa.Q, q)
    assert np.array_equal(batch.get_chain(2, sa.Q), q)
    # other chains keep Q = 1/K
    for ci in (0, 1, 3):
        assert np.all(batch.get_chain(ci, sa.Q) == 1.0 / models[ci].K)
    out = np.empty(4)
    rc = sa.lib().sgpu_get_chain_state(batch._h, len(CHAINS), sa.Q, out.ctypes.data, out.nbytes)
    assert rc == sa.ERR_ARG


def test_ragged_batch_matches_oracle_per_data_set():
    """Data sets of uneven N, L, format and K in one batch (the config-5 shape,
    scaled down): for the same (Q, P) state and stream, every chain's UpdateZ
    counts equal the oracle's exactly and its lnP(D) agrees to 1e-12."""
    import oracle
    from helpers import random_p, random_q
    rng = np.random.default_rng(55)
    specs = [("snp", 50, 300, 2), ("snp", 130, 1200, 5), ("snp", 77, 517, 8),
             ("byte", 40, 200, 3), ("byte", 90, 450, 6), ("byte", 61, 333, 4)]
    dsets, genos = [], []
    for j, (kind, N, L, _) in enumerate(specs):
        if kind == "snp":
            d = sa.synth_snp(N, L, Ktrue=3, qalpha=0.5, miss=0.03, seed=60 + j, want_geno=True)
            dsets.append((sa.GENO_SNP2, d["words"], d["nall"]))
            genos.append((d["geno"], d["nall"], 2))
        else:
            d = sa.synth_microsat(N, L, amin=2, amax=9, Ktrue=3, miss=0.03, seed=60 + j)
            fmt, words = sa.pack(d["geno"], sa.GENO_BYTE)
            dsets.append((fmt, words, d["nall"]))
            genos.append((d["geno"], d["nall"], d["A"]))
    models = [sa.make_model(N, L, K, genos[j][2], seed=70 + j, chain=j)
              for j, (_, N, L, K) in enumerate(specs)]
    batch = sa.Batch(models, dsets, list(range(len(specs))))
    batch.init()
    states = []
    for j, (_, N, L, K) in enumerate(specs):
        Q = random_q(rng, N, K, a=0.7)
        P = random_p(rng, L, K, genos[j][2], genos[j][1], a=1.0)
        batch.set_chain(j, sa.Q, Q)
        batch.set_chain(j, sa.P, P)
        states.append((Q, P))
    for rep in (5, 6):
        batch.update_z(rep, like=True)
        batch.collect(rep)
        for j, (_, N, L, K) in enumerate(specs):
            geno, nall, A = genos[j]
            och = oracle.Chain(oracle.make_model(N, L, K, A, seed=70 + j, chain=j), geno, nall)
            Q, P = states[j]
            nA, nQ = och.update_z(Q, P, rep)
            assert np.array_equal(batch.get_chain(j, sa.NA), nA), (rep, j)
            assert np.array_equal(batch.get_chain(j, sa.NQ), nQ), (rep, j)
            like = batch.get_chain(j, sa.LIKE)[0]
            ref = och.calc_like(Q, P)
            assert abs(like - ref) <= 1e-12 * abs(ref), (rep, j, like, ref)
    batch.close()
