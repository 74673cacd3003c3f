"""The largest single chain at its full size: N=20000 individuals, L=500000
biallelic loci (2-bit words, 2.5 GB), K=8, F model with alpha and lambda
inferred, Metropolis every second sweep.  It is the one shape with more than
2^32 (individual, locus) sites, so every index and size of the tile partials
(625 MB of ancestry partials) is exercised.  After initialisation and four
sweeps (two Metropolis, two Gibbs, two collecting): the counts cover every
non-missing copy exactly once per locus and per individual, the allele-1
counts per locus equal the data's, and lnP(D) of the final (Q, P) equals the
FP64 oracle's CalcLike (CalcLike.c:387-426) within 1e-9 relative.  The oracle
reads the same packed words on the host threads (its int8 form would take
20 GB)."""
import time

import numpy as np
import pytest

import oracle
import structure_amd as sa

pytestmark = pytest.mark.gpu

N, L, K = 20000, 500000, 8


@pytest.mark.timeout(900)
def test_config3_counts_and_lnp_at_full_size():
    t0 = time.perf_counter()
    print(f"config 3: generating N={N} L={L} (Balding-Nichols, K_true=8)", flush=True)
    d = sa.synth_snp(N, L, Ktrue=8, flo=0.005, fhi=0.05, qalpha=0.2, miss=0.01, seed=5)
    words, nall = d["words"], d["nall"]
    assert N * L > 2 ** 32
    print(f"config 3: data in {time.perf_counter() - t0:.1f}s; GPU chain", flush=True)
    kw = dict(freqscorr=1, inferalpha=1, inferlambda=1, metrofreq=2, burnin=2, seed=17, chain=1)
    ch = sa.Chain(sa.make_model(N, L, K, 2, **kw), sa.GENO_SNP2, words, nall)
    ch.init()
    ch.run(0, 4)
    ch.synchronize()
    nA, nQ = ch.get(sa.NA), ch.get(sa.NQ)
    Q, P, like = ch.get(sa.Q), ch.get(sa.P), ch.get(sa.LIKE)
    nsum = ch.get(sa.NSUM)[0]
    ch.close()
    print(f"config 3: 4 sweeps done at {time.perf_counter() - t0:.1f}s; host recounts",
          flush=True)
    assert nsum == 2
    lv, la, iv = oracle.snp2_counts(words, N, L)
    assert lv.sum() > 0.98 * 2 * N * L
    assert np.array_equal(nA.sum(axis=(1, 2), dtype=np.int64), lv)
    assert np.array_equal(nA[:, :, 1].sum(axis=1, dtype=np.int64), la)
    assert np.array_equal(nQ.sum(axis=1, dtype=np.int64), iv)
    assert (nA >= 0).all() and (nQ >= 0).all()
    print(f"config 3: counts OK at {time.perf_counter() - t0:.1f}s; oracle lnP(D)", flush=True)
    assert np.all(np.isfinite(Q)) and np.allclose(Q.sum(axis=1), 1.0, atol=1e-12)
    ref = oracle.calc_like_snp2(oracle.make_model(N, L, K, 2), words, Q, P)
    rel = abs(like[0] - ref) / abs(ref)
    print(f"config 3: lnP GPU {like[0]:.6f} oracle {ref:.6f} rel {rel:.2e} "
          f"({time.perf_counter() - t0:.1f}s)", flush=True)
    assert rel <= 1e-9
