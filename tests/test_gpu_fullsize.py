"""Full-size checks at the benchmark shape (N=2000, L=100000 SNPs): lnP(D) of
each chain of a K-sweep batch against the FP64 oracle's CalcLike on the same
(Q, P), within 1e-9 relative, and the size-independent invariants of the
counts (every non-missing copy counted once per locus and per individual)."""
import numpy as np
import pytest

import oracle
import structure_amd as sa

pytestmark = pytest.mark.gpu

N, L = 2000, 100000


@pytest.fixture(scope="module")
def data():
    print("full-size data N=2000 L=100000", flush=True)
    return sa.synth_snp(N, L, Ktrue=5, flo=0.01, fhi=0.1, qalpha=0.3, miss=0.01, seed=3,
                        want_geno=True)


def test_lnp_per_k_matches_fp64_oracle(data):
    Ks = (1, 2, 5, 8, 10)
    models = [sa.make_model(N, L, K, 2, freqscorr=1, inferalpha=1, inferlambda=1, metrofreq=10,
                            burnin=4, seed=11, chain=K) for K in Ks]
    b = sa.Batch(models, [(sa.GENO_SNP2, data["words"], data["nall"])])
    b.init()
    b.run(0, 12)   # 8 collecting sweeps, the last one's lnP is like3[0]
    for i, K in enumerate(Ks):
        Q, P = b.get_chain(i, sa.Q), b.get_chain(i, sa.P)
        like = b.get_chain(i, sa.LIKE)[0]
        och = oracle.Chain(oracle.make_model(N, L, K, 2), data["geno"], data["nall"])
        ref = och.calc_like(Q, P)
        print(f"K={K}: GPU lnP {like:.6f} oracle {ref:.6f} rel {abs(like - ref) / abs(ref):.2e}",
              flush=True)
        assert abs(like - ref) <= 1e-9 * abs(ref)


def test_count_invariants_at_full_size(data):
    m = sa.make_model(N, L, 6, 2, freqscorr=1, seed=5)
    ch = sa.Chain(m, sa.GENO_SNP2, data["words"], data["nall"])
    ch.init()
    ch.run(0, 3)
    nA, nQ = ch.get(sa.NA), ch.get(sa.NQ)
    g = data["geno"]
    per_locus = (g != -9).sum(axis=(0, 1))
    per_ind = (g != -9).sum(axis=(1, 2))
    assert np.array_equal(nA.sum(axis=(1, 2)), per_locus)
    assert np.array_equal(nQ.sum(axis=1), per_ind)
    # allele-1 copies per locus are fixed by the data whatever the ancestry
    assert np.array_equal(nA[:, :, 1].sum(axis=1), (g == 1).sum(axis=(0, 1)))


@pytest.mark.parametrize("fmt,N2,L2,K2", [("snp", 150000, 3000, 4), ("byte", 60000, 700, 5)])
def test_count_invariants_at_large_n(fmt, N2, L2, K2):
    """Many individuals (up to 150,000): every index of the tile partials and of
    the individual-major arrays stays in range, the sweep's counts still cover
    every non-missing copy exactly once, and lnP(D) matches the oracle."""
    print(f"large-N case {fmt} N={N2} L={L2} K={K2}", flush=True)
    if fmt == "snp":
        d = sa.synth_snp(N2, L2, Ktrue=K2, qalpha=0.5, miss=0.02, seed=9, want_geno=True)
        g, nall, A, words, f = d["geno"], d["nall"], 2, d["words"], sa.GENO_SNP2
    else:
        d = sa.synth_microsat(N2, L2, amin=2, amax=12, Ktrue=K2, miss=0.02, seed=9)
        g, nall, A = d["geno"], d["nall"], d["A"]
        f, words = sa.pack(g, sa.GENO_BYTE)
    m = sa.make_model(N2, L2, K2, A, freqscorr=1, inferalpha=1, metrofreq=2, burnin=2, seed=7)
    ch = sa.Chain(m, f, words, nall)
    ch.init()
    ch.run(0, 5)   # Metropolis and Gibbs Q updates, three collecting sweeps
    nA, nQ = ch.get(sa.NA), ch.get(sa.NQ)
    assert np.array_equal(nA.sum(axis=(1, 2)), (g != -9).sum(axis=(0, 1)))
    assert np.array_equal(nQ.sum(axis=1), (g != -9).sum(axis=(1, 2)))
    for a in range(A):
        assert np.array_equal(nA[:, :, a].sum(axis=1), (g == a).sum(axis=(0, 1)))
    Q, P = ch.get(sa.Q), ch.get(sa.P)
    och = oracle.Chain(oracle.make_model(N2, L2, K2, A), g, nall)
    ref = och.calc_like(Q, P)
    like = ch.get(sa.LIKE)[0]
    assert abs(like - ref) <= 1e-9 * abs(ref)
    ch.close()
