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
