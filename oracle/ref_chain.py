"""Whole chains of the reference's own CPU sweep (oracle/_ref) -- test
infrastructure only.

The reference's sweep (structure.c:497-638) with its CPU update functions:
UpdateP (UpdateP.c:106-199), UpdateQ (UpdateQ.c:712-747: Metropolis every
METROFREQ-th sweep, else Gibbs), UpdateZ (UpdateZ.c:13-56), UpdateAlpha
(UpdateAlpha.c:189-253), UpdateLambda / UpdatePopLambda (UpdateLambda.c),
UpdateEpsilon (UpdateEpsilon.c:148-167) and UpdateFst (UpdateFst.c:142-204),
then DataCollection (output.c:50-101) after the burn-in.  Initial state as
Initialization (Init.c:329-444): Z uniform (InitializeZ), Q = 1/K, epsilon and
F from InitFreqPriors, alpha = ALPHA, lambda = LAMBDA.  The reference's own
random streams are replaced by numpy: randomArr blocks of uniforms where the
CPU functions take them, and the rand() queue of ref_harness.c elsewhere.
Used to produce chain-level statistics (posterior-mean Q, ln P(D), alpha, F)
that the GPU chains must reproduce within Monte Carlo error.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import ref

MAXRANDOM = 64   # structuredefs.h:11


def _lib():
    L = ref.lib()
    vp = C.c_void_p
    D, I = ref._D, ref._I
    sig = {
        "UpdateQ": (None, [I, I, D, D, I, D, C.c_int, vp, D, I, D, D]),
        "UpdateZ": (None, [I, D, D, I, D]),
        "UpdateEpsilon": (None, [D, D, D, I, C.c_double]),
        "DataCollection": (None, [I, I, D, D, I, I, D, D, D, D, I, I, D, D, D, D, D, D, D, D,
                                  D, D, C.c_double, D, D, I, D, D, C.c_int, D, D, C.c_int]),
        "InitFreqPriors": (None, [D, D, I, I]),
        "InitializeZ": (None, [I, vp, I]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def run_chain(geno, nall, K, A, seed, burnin, numreps, freqscorr=1, inferalpha=1,
              inferlambda=0, popspecificlambda=0, popalphas=0, onefst=0, metrofreq=10,
              alpha=1.0, lambda_=1.0, alphapropsd=0.025, alphamax=10.0, fpriormean=0.01,
              fpriorsd=0.05, unifprioralpha=1):
    """One reference chain; returns the posterior sums and lnP statistics."""
    L_ = _lib()
    geno = np.asarray(geno)
    N, _, L = geno.shape
    ref.configure(N, L, K, A, FREQSCORR=freqscorr, INFERALPHA=inferalpha,
                  INFERLAMBDA=inferlambda, POPSPECIFICLAMBDA=popspecificlambda,
                  POPALPHAS=popalphas, ONEFST=onefst, METROFREQ=metrofreq, BURNIN=burnin,
                  NUMREPS=numreps, ALPHA=alpha, LAMBDA=lambda_, ALPHAPROPSD=alphapropsd,
                  ALPHAMAX=alphamax, FPRIORMEAN=fpriormean, FPRIORSD=fpriorsd,
                  UNIFPRIORALPHA=unifprioralpha, COMPUTEPROB=1, NOALPHA=0, LOCPRIOR=0,
                  USEPOPINFO=0, NOADMIX=0, LINKAGE=0)
    rng = np.random.default_rng(seed)
    rmax = L_.sref_rand_max()
    g32 = np.ascontiguousarray(geno, dtype=np.int32)   # GenPos layout [N][2][L], -9 missing
    nall = np.ascontiguousarray(nall, dtype=np.int32)
    inds = ref.individuals(N)
    Q = np.full((N, K), 1.0 / K)
    P = np.zeros((L, K, A))
    eps = np.zeros((L, A))
    F = np.zeros(K)
    lam = np.full(K, lambda_)
    alph = np.full(K, alpha)
    Z = np.zeros((N, 2, L), np.int32)
    QSum, PSum = np.zeros((N, K)), np.zeros((L, K, A))
    FstSum, epsSum = np.zeros(K), np.zeros((L, A))
    alphaSum, lamSum = np.zeros(K), np.zeros(K)
    like, sumlikes, sumsqlikes = np.zeros(1), np.zeros(1), np.zeros(1)
    sumind, indnorm = np.zeros(N), np.zeros(N)
    dummy = np.zeros(max(N, 8))
    idummy = np.zeros(max(N, 8), np.int32)

    def queue(n=120000):
        ref.set_rand(rng.integers(0, rmax, size=n, dtype=np.int64).astype(np.int32))

    queue()
    L_.InitFreqPriors(ref.dp(eps), ref.dp(F), ref.ip(g32), ref.ip(nall))
    if not freqscorr:
        F[:] = 0.5
    L_.InitializeZ(ref.ip(g32), inds, ref.ip(Z))
    ra_p = np.empty(L * A * K * MAXRANDOM)
    ra_q = np.empty(N * MAXRANDOM)
    ra_z = np.empty(N * L * 2)
    for rep in range(burnin + numreps):
        queue()
        rng.random(out=ra_p)
        L_.UpdateP(ref.dp(P), ref.dp(eps), ref.dp(F), ref.ip(nall), ref.ip(g32), ref.ip(Z),
                   ref.dp(lam), inds, ref.dp(ra_p))
        rng.random(out=ra_q)
        L_.UpdateQ(ref.ip(g32), None, ref.dp(Q), ref.dp(P), ref.ip(Z), ref.dp(alph), rep, inds,
                   None, None, None, ref.dp(ra_q))
        rng.random(out=ra_z)
        L_.UpdateZ(ref.ip(Z), ref.dp(Q), ref.dp(P), ref.ip(g32), ref.dp(ra_z))
        if inferalpha:
            L_.UpdateAlpha(ref.dp(Q), ref.dp(alph), inds, rep)
        if inferlambda:
            if popspecificlambda:
                L_.UpdatePopLambda(ref.dp(P), ref.dp(lam), ref.ip(nall))
            else:
                L_.UpdateLambda(ref.dp(P), ref.dp(eps), ref.dp(lam), ref.ip(nall))
        if freqscorr:
            L_.UpdateEpsilon(ref.dp(P), ref.dp(eps), ref.dp(F), ref.ip(nall), float(lam[0]))
            L_.UpdateFst(ref.dp(eps), ref.dp(F), ref.dp(P), ref.ip(nall))
        if rep + 1 > burnin:
            L_.DataCollection(ref.ip(g32), None, ref.dp(Q), ref.dp(QSum), ref.ip(Z), ref.ip(Z),
                              ref.dp(P), ref.dp(PSum), ref.dp(F), ref.dp(FstSum), ref.ip(nall),
                              ref.ip(idummy), ref.dp(alph), ref.dp(alphaSum), ref.dp(dummy),
                              ref.dp(dummy), ref.dp(like), ref.dp(sumlikes),
                              ref.dp(sumsqlikes), ref.dp(dummy), ref.dp(eps), ref.dp(epsSum),
                              0.0, ref.dp(lam), ref.dp(lamSum), None, ref.dp(dummy),
                              ref.dp(dummy), 0, ref.dp(sumind), ref.dp(indnorm), rep)
        if L_.sref_rand_overrun():
            raise RuntimeError("rand() queue too short for one sweep")
    n = numreps
    mean = sumlikes[0] / n
    var = (sumsqlikes[0] - sumlikes[0] ** 2 / n) / (n - 1)
    return dict(qbar=QSum / n, pbar=PSum / n, alpha=alphaSum / n, fst=FstSum / n,
                lam=lamSum / n, lnp=mean - var / 2, meanlike=mean)
