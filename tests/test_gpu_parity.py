"""GPU parity: every HIP phase against the CPU oracle on identical state and
the same Philox stream, through the C-ABI (libsgpu.so).

Bars: Z, allele/ancestry counts and accept/reject decisions bit-exact; FP64
draws (P, Q, alpha, lambda, eps, F) within 1e-12 relative (device and glibc
transcendentals may differ in the last ulp); lnP(D) within 1e-12 relative.
"""
import numpy as np
import pytest

import oracle
import structure_amd as sa
from helpers import random_geno, random_p, random_q

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _pair(geno, nall, K, A, fmt=None, **kw):
    N, _, L = geno.shape
    om = oracle.make_model(N, L, K, A, **kw)
    gm = sa.make_model(N, L, K, A, **kw)
    och = oracle.Chain(om, geno, nall)
    gch = sa.Chain.from_geno(gm, geno, nall, fmt)
    return och, gch


def _snp(N, L, K, seed, miss=0.02):
    d = sa.synth_snp(N, L, Ktrue=K, qalpha=0.5, miss=miss, seed=seed, want_geno=True)
    return d["geno"], d["nall"], 2


def _close(a, b, rtol=RTOL):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape
    err = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    assert np.all((a == b) | (err <= rtol)), f"max rel err {err.max():.3e}"


def _state(gch, nall, rng, K, A, fq=1.0):
    m = gch.model
    Q = random_q(rng, m.N, K, a=fq)
    P = random_p(rng, m.L, K, A, nall, a=1.0)
    gch.set(sa.Q, Q)
    gch.set(sa.P, P)
    return Q, P


CASES_Z = [("snp", 1), ("snp", 2), ("snp", 5), ("snp", 10), ("byte", 3), ("byte", 7),
           ("byteraw", 4), ("byteraw", 12),
           # MAXPOPS above 16, up to 20 (UpdateZ.c:41-46)
           ("snp", 17), ("snp", 20), ("byte", 17), ("byte", 20), ("byteraw", 19)]


@pytest.mark.parametrize("kind,K", CASES_Z)
def test_update_z_counts_and_like_bitexact(kind, K):
    rng = np.random.default_rng(1000 + K)
    if kind == "snp":
        geno, nall, A = _snp(77, 600, min(K, 5), seed=K)
        fmt = sa.GENO_SNP2
    elif kind == "byte":
        d = sa.synth_microsat(45, 333, amin=2, amax=9, Ktrue=3, seed=K)
        geno, nall, A = d["geno"], d["nall"], d["A"]
        fmt = sa.GENO_BYTE
    else:   # per-copy missingness and heterozygotes in both orders
        geno, nall = random_geno(rng, 39, 290, 6, miss=0.1)
        A, fmt = 6, sa.GENO_BYTE
    och, gch = _pair(geno, nall, K, A, fmt=fmt, seed=77, chain=5)
    assert gch.fmt == fmt
    gch.init()
    Q, P = _state(gch, nall, rng, K, A, fq=0.6)
    for rep in (3, 4):
        z = gch.update_z(rep, like=True, want_z=True)
        nA, nQ, Zo = och.update_z(Q, P, rep, want_z=True)
        assert np.array_equal(z.astype(np.int32), Zo)
        assert np.array_equal(gch.get(sa.NA), nA)
        assert np.array_equal(gch.get(sa.NQ), nQ)
        gch.collect(rep)
        like = gch.get(sa.LIKE)[0]
        _close(like, och.calc_like(Q, P))


@pytest.mark.parametrize("N", [30, 100, 200, 300])
def test_update_p_sums_every_tile_partial(N):
    """UpdateP on the per-tile allele-count partials that UpdateZ itself leaves
    (a host write of the counts fills only the first tile).  With few loci each
    16-individual word row is a tile of its own, so N = 30, 100, 200 and 300
    give 2, 7, 13 and 19 partials: the straight-line sums of 4, 8 and 16 loads
    and the loop beyond 16.  Two-allele and F-model parameters, P against the
    oracle's UpdateP on the oracle's counts."""
    rng = np.random.default_rng(3000 + N)
    geno, nall, A = _snp(N, 300, 3, seed=N)
    K = 3
    och, gch = _pair(geno, nall, K, A, fmt=sa.GENO_SNP2, seed=31, chain=1, freqscorr=1)
    gch.init()
    och.init()
    Q, P = _state(gch, nall, rng, K, A, fq=0.6)
    eps = och.eps.copy()
    F = rng.uniform(0.01, 0.4, K)
    gch.set(sa.FST, F)
    lam = np.ones(K)
    for rep in (4, 5):
        gch.update_z(rep, like=False)
        nA, _ = och.update_z(Q, P, rep)
        assert np.array_equal(gch.get(sa.NA), nA)
        gch.update_p(rep, collect=False)
        P = och.update_p(nA, eps, F, lam, rep)
        _close(gch.get(sa.P), P)
        gch.set(sa.P, P)


@pytest.mark.parametrize("K", list(range(1, 21)))
def test_burnin_update_z_counts_bitexact(K):
    """UpdateZ without the lnP terms (burn-in sweeps), 2-bit words, every K up to
    20: the variant forms each copy's scaled uniform against both alleles (four
    products) at most K and keeps the selects at K = 4, 7, 10, 11 and 14-16; the
    counts equal the oracle's either way."""
    rng = np.random.default_rng(2000 + K)
    geno, nall, A = _snp(70, 520, min(K, 5), seed=40 + K)
    och, gch = _pair(geno, nall, K, A, fmt=sa.GENO_SNP2, seed=78, chain=4)
    gch.init()
    Q, P = _state(gch, nall, rng, K, A, fq=0.6)
    for rep in (5, 6):
        gch.update_z(rep, like=False)
        nA, nQ = och.update_z(Q, P, rep)
        assert np.array_equal(gch.get(sa.NA), nA), (K, rep)
        assert np.array_equal(gch.get(sa.NQ), nQ), (K, rep)


@pytest.mark.parametrize("kind", ["snp", "byte"])
def test_init_state_matches_oracle(kind):
    if kind == "snp":
        geno, nall, A = _snp(50, 700, 3, seed=11)
    else:
        d = sa.synth_microsat(40, 120, amin=2, amax=8, Ktrue=3, seed=11)
        geno, nall, A = d["geno"], d["nall"], d["A"]
    for freqscorr in (1, 0):
        och, gch = _pair(geno, nall, 4, A, seed=9, chain=1, freqscorr=freqscorr, lambda0=0.7)
        gch.init()
        och.init()
        assert np.array_equal(gch.get(sa.NA), och.nA)
        assert np.array_equal(gch.get(sa.NQ), och.nQ)
        _close(gch.get(sa.EPS), och.eps, 1e-15)
        assert np.array_equal(gch.get(sa.FST), och.F)
        assert np.array_equal(gch.get(sa.Q), och.Q)
        assert np.array_equal(gch.get(sa.ALPHA), och.alpha)


@pytest.mark.parametrize("kind", ["microsat", "biallelic"])
@pytest.mark.parametrize("freqscorr", [1, 0])
def test_update_p_matches_oracle(freqscorr, kind):
    rng = np.random.default_rng(7)
    if kind == "microsat":
        d = sa.synth_microsat(30, 200, amin=1, amax=12, Ktrue=3, seed=5)
        geno, nall, A = d["geno"], d["nall"], d["A"]
    else:   # A = 2 (one thread per locus and pop), with monomorphic and empty loci
        geno, nall = random_geno(rng, 30, 300, 2, miss=0.05)
        nall[::7] = 1
        geno[:, :, ::7] = np.where(geno[:, :, ::7] >= 0, 0, -9)
        nall[::11] = 0
        geno[:, :, ::11] = -9
        A = 2
    K = 4
    och, gch = _pair(geno, nall, K, A, seed=3, chain=2, freqscorr=freqscorr)
    gch.init()
    och.init()
    Q, P = _state(gch, nall, rng, K, A)
    nA, _ = och.update_z(Q, P, 1)
    gch.set(sa.NA, nA)
    eps = och.eps.copy()
    F = rng.uniform(0.01, 0.4, K)
    lam = rng.uniform(0.2, 2.0, K)
    gch.set(sa.FST, F)
    gch.set(sa.LAMBDA, lam)
    for rep in (2, 3):
        gch.update_p(rep, collect=True)
        _close(gch.get(sa.P), och.update_p(nA, eps, F, lam, rep))
    psum = gch.get(sa.PSUM)
    _close(psum, och.update_p(nA, eps, F, lam, 2) + och.update_p(nA, eps, F, lam, 3))


def test_update_q_gibbs_and_metro_match_oracle():
    rng = np.random.default_rng(8)
    geno, nall, A = _snp(70, 500, 3, seed=21)
    K = 3
    och, gch = _pair(geno, nall, K, A, seed=4, chain=0, metrofreq=10)
    gch.init()
    och.init()
    Q, P = _state(gch, nall, rng, K, A)
    _, nQ = och.update_z(Q, P, 1)
    gch.set(sa.NQ, nQ)
    alpha = np.array([0.4, 1.3, 0.8])
    gch.set(sa.ALPHA, alpha)
    gch.update_q(7)
    _close(gch.get(sa.Q), och.update_q_gibbs(nQ, alpha, 7))
    nacc = 0
    for rep in (10, 20, 30):
        Q0 = random_q(rng, 70, K, a=1.0)
        gch.set(sa.Q, Q0)
        gch.update_q(rep)
        Qo, acc = och.update_q_metro(P, alpha, rep, Q0)
        _close(gch.get(sa.Q), Qo)
        nacc += acc.sum()
    assert 0 < nacc < 3 * 70


@pytest.mark.parametrize("popalphas,unif", [(0, 1), (1, 1), (0, 0)])
def test_update_alpha_matches_oracle(popalphas, unif):
    rng = np.random.default_rng(9)
    geno, nall, A = _snp(40, 300, 3, seed=2)
    K = 4
    och, gch = _pair(geno, nall, K, A, popalphas=popalphas, unifprioralpha=unif,
                     alphapropsd=0.2, alphamax=2.5, seed=12)
    gch.init()
    moved = 0
    for rep in range(40):
        Q = random_q(rng, 40, K, a=rng.uniform(0.3, 2))
        alpha = rng.uniform(0.2, 2.0, K) if popalphas else np.full(K, rng.uniform(0.2, 2))
        gch.set(sa.Q, Q)
        gch.set(sa.ALPHA, alpha)
        gch.update_alpha(rep)
        ao = och.update_alpha(Q, alpha, rep)
        _close(gch.get(sa.ALPHA), ao)
        moved += int(not np.array_equal(ao, alpha))
    assert 3 < moved < 40


@pytest.mark.parametrize("freqscorr,popspec", [(1, 0), (0, 0), (0, 1)])
def test_update_lambda_matches_oracle(freqscorr, popspec):
    rng = np.random.default_rng(10)
    d = sa.synth_microsat(30, 150, amin=1, amax=6, Ktrue=3, seed=9)
    geno, nall, A = d["geno"], d["nall"], d["A"]
    K = 3
    och, gch = _pair(geno, nall, K, A, freqscorr=freqscorr, popspecificlambda=popspec,
                     inferlambda=1, seed=13)
    gch.init()
    och.init()
    moved = 0
    for rep in range(30):
        P = random_p(rng, 150, K, A, nall, a=rng.uniform(0.3, 3))
        eps = np.zeros((150, A))
        for l in range(150):
            if nall[l]:
                eps[l, :nall[l]] = rng.dirichlet(np.ones(nall[l]))
        lam = rng.uniform(0.2, 3, K) if popspec else np.full(K, rng.uniform(0.2, 3))
        gch.set(sa.P, P)
        gch.set(sa.EPS, eps)
        gch.set(sa.LAMBDA, lam)
        gch.update_lambda(rep)
        lo = och.update_lambda(P, eps, lam, rep)
        _close(gch.get(sa.LAMBDA), lo)
        moved += int(not np.array_equal(lo, lam))
    assert 2 < moved < 30


@pytest.mark.parametrize("kind", ["microsat", "biallelic"])
@pytest.mark.parametrize("lam0", [1.0, 0.6])
def test_update_eps_matches_oracle(lam0, kind):
    """Both the general path and the two-allele one (A = 2: both alleles' eps and
    frequencies loaded before the draw picks a1), with monomorphic and empty
    loci among the two-allele ones."""
    rng = np.random.default_rng(11)
    if kind == "microsat":
        d = sa.synth_microsat(25, 400, amin=1, amax=7, Ktrue=3, seed=10)
        geno, nall, A = d["geno"], d["nall"], d["A"]
    else:
        geno, nall = random_geno(rng, 25, 400, 2, miss=0.05)
        nall[::7] = 1
        geno[:, :, ::7] = np.where(geno[:, :, ::7] >= 0, 0, -9)
        nall[::11] = 0
        geno[:, :, ::11] = -9
        A = 2
    K = 3
    och, gch = _pair(geno, nall, K, A, seed=14)
    gch.init()
    och.init()
    P = random_p(rng, 400, K, A, nall, a=2.0)
    F = rng.uniform(0.02, 0.3, K)
    eps = och.eps.copy()
    gch.set(sa.P, P)
    gch.set(sa.FST, F)
    gch.set(sa.LAMBDA, np.full(K, lam0))
    gch.set(sa.EPS, eps)
    for rep in (1, 2, 3):
        gch.update_eps(rep, collect=True)
        eps = och.update_eps(P, F, lam0, eps, rep)
        _close(gch.get(sa.EPS), eps)
    assert not np.array_equal(eps, och.eps)


@pytest.mark.parametrize("onefst", [0, 1])
def test_update_fst_matches_oracle(onefst):
    rng = np.random.default_rng(12)
    d = sa.synth_microsat(25, 300, amin=1, amax=5, Ktrue=3, seed=12)
    geno, nall, A = d["geno"], d["nall"], d["A"]
    K = 3
    och, gch = _pair(geno, nall, K, A, onefst=onefst, fpriormean=0.05, fpriorsd=0.05, seed=15)
    gch.init()
    och.init()
    moved = 0
    for rep in range(30):
        P = random_p(rng, 300, K, A, nall, a=2.0)
        eps = och.eps.copy()
        F = np.full(K, rng.uniform(0.02, 0.2)) if onefst else rng.uniform(0.02, 0.2, K)
        gch.set(sa.P, P)
        gch.set(sa.EPS, eps)
        gch.set(sa.FST, F)
        gch.update_fst(rep)
        fo = och.update_fst(eps, P, F, rep)
        _close(gch.get(sa.FST), fo)
        moved += int(not np.array_equal(fo, F))
    assert 2 < moved < 30


@pytest.mark.parametrize("kind", ["microsat", "snp_fmodel_lambda"])
def test_chain_trajectory_matches_oracle(kind):
    """Whole sweeps: the GPU chain and the oracle chain follow the same path."""
    if kind == "microsat":   # config-1 shape, scaled down
        d = sa.synth_microsat(60, 90, amin=2, amax=10, Ktrue=3, qalpha=0.5, miss=0.02, seed=1)
        geno, nall, A = d["geno"], d["nall"], d["A"]
        kw = dict(freqscorr=0, inferalpha=1, lambda0=1.0, burnin=8, metrofreq=10)
        K = 3
    else:                    # config-2 options, small shape
        geno, nall, A = _snp(96, 400, 5, seed=3, miss=0.01)
        kw = dict(freqscorr=1, inferalpha=1, inferlambda=1, burnin=8, metrofreq=10)
        K = 5
    och, gch = _pair(geno, nall, K, A, seed=2024, chain=1, **kw)
    gch.init()
    och.init()
    for rep in range(31):
        gch.sweep(rep)
        och.sweep(rep)
    for sel, ref in [(sa.Q, och.Q), (sa.P, och.P), (sa.EPS, och.eps), (sa.FST, och.F),
                     (sa.ALPHA, och.alpha), (sa.LAMBDA, och.lambda_), (sa.QSUM, och.QSum),
                     (sa.PSUM, och.PSum), (sa.ALPHASUM, och.alphaSum),
                     (sa.FSTSUM, och.FstSum), (sa.EPSSUM, och.epsSum)]:
        _close(gch.get(sel), ref, 1e-9)
    assert np.array_equal(gch.get(sa.NA), och.nA)
    assert np.array_equal(gch.get(sa.NQ), och.nQ)
    like = gch.get(sa.LIKE)
    _close(like, [och.like, och.sumlikes, och.sumsqlikes], 1e-9)
    assert gch.get(sa.NSUM)[0] == och.state.nsum == 23


TINY = [  # (format, N, L, K): single individuals and loci, odd N, L off every tile size
    ("snp", 1, 1, 1), ("snp", 1, 5, 2), ("snp", 3, 1, 3), ("snp", 17, 3, 2),
    ("snp", 33, 257, 4), ("snp", 2, 300, 16),
    ("byte", 1, 1, 2), ("byte", 3, 7, 3), ("byte", 5, 129, 5), ("byte", 2, 33, 16),
    # K above 16: the largest register tiles, 128-locus Metropolis staging (2-bit)
    ("snp", 40, 300, 17), ("snp", 33, 257, 20), ("byte", 9, 40, 17), ("byte", 5, 129, 20),
]


def _follow_oracle(fmt, N, L, K, A_byte=5, sweeps=7):
    """Whole sweeps (Gibbs and Metropolis Q, F model, alpha and lambda) with
    graph replay on a small shape: the same trajectory as the oracle."""
    rng = np.random.default_rng(N * 1000 + L * 10 + K + A_byte)
    if fmt == "snp":   # 2-bit words: biallelic, both copies missing together
        geno, nall = random_geno(rng, N, L, 2, miss=0.0, biallelic=True)
        geno = np.sort(geno, axis=1)   # heterozygotes as (0, 1)
        gone = rng.random((N, L)) < 0.1
        for c in range(2):
            geno[:, c, :] = np.where(gone, -9, geno[:, c, :])
        A, f = 2, sa.GENO_SNP2
    else:
        geno, nall = random_geno(rng, N, L, A_byte, miss=0.1)
        A, f = A_byte, sa.GENO_BYTE
    kw = dict(freqscorr=1, inferalpha=1, inferlambda=1, burnin=3, metrofreq=2)
    och, gch = _pair(geno, nall, K, A, fmt=f, seed=99, chain=K, **kw)
    assert gch.fmt == f
    gch.init()
    och.init()
    gch.run(0, sweeps)
    for rep in range(sweeps):
        och.sweep(rep)
    assert np.array_equal(gch.get(sa.NA), och.nA)
    assert np.array_equal(gch.get(sa.NQ), och.nQ)
    for sel, ref in [(sa.Q, och.Q), (sa.P, och.P), (sa.EPS, och.eps), (sa.FST, och.F),
                     (sa.ALPHA, och.alpha), (sa.LAMBDA, och.lambda_), (sa.QSUM, och.QSum),
                     (sa.PSUM, och.PSum)]:
        _close(gch.get(sel), ref, 1e-9)
    _close(gch.get(sa.LIKE), [och.like, och.sumlikes, och.sumsqlikes], 1e-9)


@pytest.mark.parametrize("fmt,N,L,K", TINY)
def test_tiny_and_odd_shapes_follow_the_oracle(fmt, N, L, K):
    _follow_oracle(fmt, N, L, K)


@pytest.mark.parametrize("N,L,K,A", [(9, 11, 16, 120), (40, 6, 3, 127), (130, 3, 8, 64),
                                     (9, 11, 20, 127), (12, 9, 4, 200), (7, 5, 20, 255),
                                     (70, 4, 3, 255)])
def test_many_alleles_follow_the_oracle(N, L, K, A):
    """MAXALLELES up to 255 (the byte format's largest: alleles 0..254; above 127
    packed from wider integers by sgpu_pack_byte32): one locus per LDS chunk in
    UpdateZ (61 KiB at K = 20, A = 255), the Metropolis ratio's staging sized to
    one locus, one (locus, pop) group per UpdateP block."""
    _follow_oracle("byte", N, L, K, A_byte=A)


@pytest.mark.parametrize("variant", ["full", "no_tail_in_burnin"])
def test_graph_replays_match_direct_launches(variant):
    """sgpu_run_sweeps replays captured HIP graphs (one per sweep type) with the
    sweep index read on the device; every replay block must equal the same
    sweeps launched one by one, with host reads and syncs in between.  In the
    second variant (no F model, alpha fixed) burn-in sweeps have no closing
    single-block launch, so the index advance is a launch of its own there."""
    geno, nall, A = _snp(80, 700, 4, seed=8, miss=0.02)
    if variant == "full":
        kw = dict(freqscorr=1, inferalpha=1, inferlambda=1, burnin=9, metrofreq=5, seed=31)
    else:
        kw = dict(freqscorr=0, inferalpha=0, inferlambda=1, burnin=9, metrofreq=5, seed=31)
    ga = sa.Chain.from_geno(sa.make_model(80, 700, 4, A, **kw), geno, nall)
    gb = sa.Chain.from_geno(sa.make_model(80, 700, 4, A, **kw), geno, nall)
    ga.init()
    gb.init()
    rep = 0
    for block in (7, 7, 6, 11):
        ga.run(rep, block)
        for r in range(rep, rep + block):
            gb.sweep(r)
        rep += block
        for sel in (sa.Q, sa.P, sa.EPS, sa.FST, sa.ALPHA, sa.LAMBDA, sa.NA, sa.NQ, sa.QSUM,
                    sa.PSUM, sa.LIKE, sa.NSUM):
            assert np.array_equal(ga.get(sel), gb.get(sel)), (block, sel)
