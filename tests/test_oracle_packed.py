"""The oracle's 2-bit-word helpers (oracle/oracle_packed.c), used by the
full-size checks, against the int8 oracle and numpy recounts: CalcLike over
packed words with the individuals split over threads equals the oracle's
CalcLike (itself pinned bit-exact against the reference's CalcLike) bit for
bit, and the packed recounts equal numpy counts of the unpacked genotypes."""
import numpy as np
import pytest

import oracle
from helpers import pack_snp2_words, random_p, random_q, random_snp_codes


@pytest.mark.parametrize("N,L,K,threads", [(1, 1, 1, 1), (17, 300, 3, 2), (50, 513, 5, 3),
                                           (33, 257, 8, 0), (16, 1000, 20, 4)])
def test_calc_like_snp2_equals_int8_oracle(N, L, K, threads):
    rng = np.random.default_rng(N * 1000 + L)
    codes, geno = random_snp_codes(rng, N, L, miss=0.1)
    nall = np.full(L, 2, np.int32)
    Q = random_q(rng, N, K, 0.5)
    P = random_p(rng, L, K, 2, nall, 0.3)
    P[0, 0, :] = [1e-60, 1.0]   # a term under the sqrt(UNDERFLO) clamp
    m = oracle.make_model(N, L, K, 2)
    ref = oracle.Chain(m, geno, nall).calc_like(Q, P)
    words = pack_snp2_words(codes)
    got = oracle.calc_like_snp2(m, words, Q, P, threads=threads)
    assert got == ref


def test_calc_like_snp2_long_individual_renormalises():
    """Many loci per individual: the running product drops below sqrt(UNDERFLO)
    repeatedly and is logged and reset, as in CalcLikeInd."""
    rng = np.random.default_rng(5)
    N, L, K = 20, 4000, 4
    codes, geno = random_snp_codes(rng, N, L, miss=0.02)
    nall = np.full(L, 2, np.int32)
    Q = random_q(rng, N, K)
    P = random_p(rng, L, K, 2, nall, 0.2)
    m = oracle.make_model(N, L, K, 2)
    ref = oracle.Chain(m, geno, nall).calc_like(Q, P)
    got = oracle.calc_like_snp2(m, pack_snp2_words(codes), Q, P, threads=3)
    assert got == ref


@pytest.mark.parametrize("N,L,threads", [(1, 5, 1), (31, 700, 3), (64, 256, 0)])
def test_snp2_counts_equal_numpy(N, L, threads):
    rng = np.random.default_rng(N + L)
    codes, geno = random_snp_codes(rng, N, L, miss=0.07)
    lv, la, iv = oracle.snp2_counts(pack_snp2_words(codes), N, L, threads=threads)
    assert np.array_equal(lv, (geno != -9).sum(axis=(0, 1)))
    assert np.array_equal(la, (geno == 1).sum(axis=(0, 1)))
    assert np.array_equal(iv, (geno != -9).sum(axis=(1, 2)))


def test_calc_like_snp2_detects_a_dropped_term():
    """A planted change of one frequency moves the packed result: the check has power."""
    rng = np.random.default_rng(9)
    N, L, K = 40, 300, 3
    codes, geno = random_snp_codes(rng, N, L)
    nall = np.full(L, 2, np.int32)
    Q = random_q(rng, N, K)
    P = random_p(rng, L, K, 2, nall)
    m = oracle.make_model(N, L, K, 2)
    ref = oracle.Chain(m, geno, nall).calc_like(Q, P)
    P2 = P.copy()
    P2[7, 2, :] = P2[7, 2, ::-1]
    assert oracle.calc_like_snp2(m, pack_snp2_words(codes), Q, P2) != ref
