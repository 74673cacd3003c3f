"""CPU oracle for the STRUCTURE admixture sweep -- test infrastructure only.

Python (ctypes + numpy) front end of ``oracle/oracle.c``, the FP64 C
restatement of the reference CPU update functions driven by the same
Philox4x32-10 stream as the HIP kernels.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import
this package, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PH_Z, PH_INITZ, PH_P, PH_Q, PH_QMETRO, PH_ALPHA, PH_LAMBDA, PH_EPS, PH_FST = range(1, 10)
MISSING = -9


def build(force: bool = False) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    cmd = ["make", "-s", "-C", _HERE, "all"]
    if force:
        subprocess.run(["make", "-s", "-C", _HERE, "clean"], check=True)
    subprocess.run(cmd, check=True)


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", _HERE, "oracle"], check=True)
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


class Model(C.Structure):
    """Mirror of ``orc_model`` (oracle.h)."""

    _fields_ = [(n, C.c_int) for n in (
        "N", "L", "K", "A", "burnin", "numreps", "freqscorr", "onefst", "inferalpha",
        "popalphas", "unifprioralpha", "inferlambda", "popspecificlambda", "metrofreq",
        "computeprob")] + [(n, C.c_double) for n in (
            "alpha0", "lambda0", "alphapropsd", "alphamax", "alphapriora", "alphapriorb",
            "fpriormean", "fpriorsd")] + [("seed", C.c_uint32), ("chain", C.c_uint32)] + [
        ("ancestdist", C.c_int), ("numboxes", C.c_int), ("pfrompopflagonly", C.c_int),
        ("popflag", C.POINTER(C.c_int32))]


def make_model(N, L, K, A, **kw) -> Model:
    """Model with the defaults of the shipped extraparams (extraparams:1-120)."""
    d = dict(burnin=0, numreps=0, freqscorr=1, onefst=0, inferalpha=1, popalphas=0,
             unifprioralpha=1, inferlambda=0, popspecificlambda=0, metrofreq=10,
             computeprob=1, alpha0=1.0, lambda0=1.0, alphapropsd=0.025, alphamax=10.0,
             alphapriora=1.0, alphapriorb=2.0, fpriormean=0.01, fpriorsd=0.05, seed=1,
             chain=0, ancestdist=0, numboxes=1000, pfrompopflagonly=0)
    d.update(kw)
    return Model(N=N, L=L, K=K, A=A, **d)


_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_B = C.POINTER(C.c_int16)   # genotypes [N][2][L] int16 (-9 missing)
_U = C.POINTER(C.c_uint32)
_M = C.POINTER(Model)


class State(C.Structure):
    _fields_ = [(n, _D) for n in ("Q", "P", "eps", "F", "alpha", "lambda_")] + [
        ("nA", _I), ("nQ", _I)] + [(n, _D) for n in (
            "QSum", "PSum", "FstSum", "epsSum", "alphaSum", "lambdaSum")] + [
        ("like", C.c_double), ("sumlikes", C.c_double), ("sumsqlikes", C.c_double),
        ("nsum", C.c_int64), ("ancestDist", _I)]


def _declare(L: C.CDLL) -> None:
    u32 = C.c_uint32
    sig = {
        "orc_philox4x32_10": (None, [_U, _U, _U]),
        "orc_philox_block": (None, [u32, u32, u32, u32, u32, u32, _U]),
        "orc_uz": (C.c_double, [u32]),
        "orc_uo": (C.c_double, [u32]),
        "orc_normal": (C.c_double, [u32, u32, u32, u32, u32, u32]),
        "orc_gamma": (C.c_double, [C.c_double, u32, u32, u32, u32, u32, u32]),
        "orc_lgamma": (C.c_double, [C.c_double]),
        "orc_init_state": (None, [_M, _B, _I, C.POINTER(State)]),
        "orc_count_from_z": (None, [_M, _B, _I, _I, _I]),
        "orc_update_z": (None, [_M, _B, _D, _D, u32, _I, _I, _I]),
        "orc_init_z": (None, [_M, _B, _I, _I, _I]),
        "orc_update_p": (None, [_M, _I, _I, _D, _D, _D, u32, _D]),
        "orc_update_q_gibbs": (None, [_M, _I, _D, u32, _D]),
        "orc_calc_like_ind_diff": (C.c_double, [_M, _B, _D, _D, _D, C.c_int]),
        "orc_update_q_metro": (None, [_M, _B, _D, _D, u32, _D, _I]),
        "orc_log_prob_q": (C.c_double, [_M, _D, C.c_double]),
        "orc_log_prob_q_onepop": (C.c_double, [_M, _D, C.c_double, C.c_double, C.c_int]),
        "orc_alpha_prior_diff": (C.c_double, [_M, C.c_double, C.c_double]),
        "orc_update_alpha_given": (None, [_M, _D, _D, _D, _D]),
        "orc_update_alpha": (None, [_M, _D, u32, _D]),
        "orc_update_lambda_given": (None, [_M, _I, _D, _D, _D, _D, _D]),
        "orc_update_lambda": (None, [_M, _I, _D, _D, u32, _D]),
        "orc_update_eps_given": (None, [_M, _I, _D, _D, C.c_double, _D, _I, _I, _D, _D]),
        "orc_update_eps": (None, [_M, _I, _D, _D, C.c_double, u32, _D]),
        "orc_fprior_diff": (C.c_double, [_M, C.c_double, C.c_double]),
        "orc_flike_freqs": (C.c_double, [_M, _I, C.c_double, _D, _D, C.c_int]),
        "orc_update_fst_given": (None, [_M, _I, _D, _D, _D, _D, _D]),
        "orc_update_fst": (None, [_M, _I, _D, _D, u32, _D]),
        "orc_calc_like": (C.c_double, [_M, _B, _D, _D]),
        "orc_collect": (None, [_M, _B, _I, C.POINTER(State)]),
        "orc_sweep": (None, [_M, _B, _I, u32, C.POINTER(State)]),
        "orc_calc_like_snp2": (C.c_double, [_M, _U, C.c_int, _D, _D, C.c_int]),
        "orc_snp2_counts": (None, [C.c_int, C.c_int, C.c_int, _U, C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def ptr(a: np.ndarray, ct):
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ct))


def dp(a):
    assert a.dtype == np.float64
    return ptr(a, C.c_double)


def ip(a):
    assert a.dtype == np.int32
    return ptr(a, C.c_int32)


def bp(a):
    assert a.dtype == np.int16
    return ptr(a, C.c_int16)


# ---- scalar helpers ---------------------------------------------------------

def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, dtype=np.uint32).copy()
    k = np.asarray(key, dtype=np.uint32).copy()
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(ptr(c, C.c_uint32), ptr(k, C.c_uint32), ptr(out, C.c_uint32))
    return out


def philox_block(seed, chain, c0, c1, c2, c3) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox_block(seed, chain, c0, c1, c2, c3, ptr(out, C.c_uint32))
    return out


def lgamma(z: float) -> float:
    return lib().orc_lgamma(float(z))


def gamma(a, seed, chain, c0, c1, c2, phase) -> float:
    return lib().orc_gamma(float(a), seed, chain, c0, c1, c2, phase)


def normal(seed, chain, c0, c1, c2, phase) -> float:
    return lib().orc_normal(seed, chain, c0, c1, c2, phase)


# ---- 2-bit genotype words (large data, host threads) ---------------------------

def _lpad(words: np.ndarray, N: int) -> int:
    rows = (N + 15) // 16
    assert words.size % rows == 0, "words must be [ceil(N/16)][Lpad]"
    return words.size // rows


def calc_like_snp2(model: Model, words: np.ndarray, Q, P, threads=0) -> float:
    """CalcLike (CalcLike.c:387-426) of (Q, P) on 2-bit words [ceil(N/16)][Lpad],
    equal bit for bit to Chain.calc_like on the unpacked genotypes."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    P = np.ascontiguousarray(P, dtype=np.float64)
    assert Q.shape == (model.N, model.K) and P.shape == (model.L, model.K, model.A)
    return lib().orc_calc_like_snp2(C.byref(model), ptr(w, C.c_uint32), _lpad(w, model.N),
                                    dp(Q), dp(P), int(threads))


def snp2_counts(words: np.ndarray, N: int, L: int, threads=0):
    """(non-missing copies per locus, allele-1 copies per locus, non-missing
    copies per individual) of 2-bit words [ceil(N/16)][Lpad], as int64."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    lv = np.zeros(L, np.int64)
    la = np.zeros(L, np.int64)
    iv = np.zeros(N, np.int64)
    lib().orc_snp2_counts(N, L, _lpad(w, N), ptr(w, C.c_uint32), ptr(lv, C.c_int64),
                          ptr(la, C.c_int64), ptr(iv, C.c_int64), int(threads))
    return lv, la, iv


# ---- chain ------------------------------------------------------------------

@dataclass
class Chain:
    """One oracle chain: model, genotypes ([N][2][L], -9 missing; int16 inside,
    so alleles up to 254 fit), state;
    popflag ([N], the data file's POPFLAG column) for PFROMPOPFLAGONLY."""

    model: Model
    geno: np.ndarray
    nall: np.ndarray
    popflag: np.ndarray | None = None
    arrays: dict = field(default_factory=dict)

    def __post_init__(self):
        m = self.model
        N, L, K, A = m.N, m.L, m.K, m.A
        self.geno = np.ascontiguousarray(self.geno, dtype=np.int16)
        assert self.geno.shape == (N, 2, L)
        self.nall = np.ascontiguousarray(self.nall, dtype=np.int32)
        if self.popflag is not None:
            self.popflag = np.ascontiguousarray(self.popflag, dtype=np.int32)
            assert self.popflag.shape == (N,)
            m.popflag = ip(self.popflag)
        z = np.zeros
        self.arrays = dict(
            Q=z((N, K)), P=z((L, K, A)), eps=z((L, A)), F=z(K), alpha=z(K), lambda_=z(K),
            nA=z((L, K, A), np.int32), nQ=z((N, K), np.int32), QSum=z((N, K)),
            PSum=z((L, K, A)), FstSum=z(K), epsSum=z((L, A)), alphaSum=z(K),
            lambdaSum=z(K))
        if m.ancestdist:
            self.arrays["ancestDist"] = z((N, K, m.numboxes), np.int32)
        st = State()
        for name, arr in self.arrays.items():
            setattr(st, name, dp(arr) if arr.dtype == np.float64 else ip(arr))
        self.state = st

    def __getattr__(self, name):
        arrays = self.__dict__.get("arrays")
        if arrays and name in arrays:
            return arrays[name]
        raise AttributeError(name)

    def init(self):
        lib().orc_init_state(C.byref(self.model), bp(self.geno), ip(self.nall),
                             C.byref(self.state))

    def sweep(self, rep: int):
        lib().orc_sweep(C.byref(self.model), bp(self.geno), ip(self.nall), rep,
                        C.byref(self.state))

    @property
    def like(self):
        return self.state.like

    @property
    def sumlikes(self):
        return self.state.sumlikes

    @property
    def sumsqlikes(self):
        return self.state.sumsqlikes

    # per-phase entry points on explicit arrays
    def update_z(self, Q, P, rep, want_z=False):
        m = self.model
        nA = np.zeros((m.L, m.K, m.A), np.int32)
        nQ = np.zeros((m.N, m.K), np.int32)
        Z = np.zeros((m.N, 2, m.L), np.int32) if want_z else None
        lib().orc_update_z(C.byref(m), bp(self.geno), dp(Q), dp(P), rep,
                           ip(Z) if want_z else None, ip(nA), ip(nQ))
        return (nA, nQ, Z) if want_z else (nA, nQ)

    def init_z(self, want_z=False):
        m = self.model
        nA = np.zeros((m.L, m.K, m.A), np.int32)
        nQ = np.zeros((m.N, m.K), np.int32)
        Z = np.zeros((m.N, 2, m.L), np.int32) if want_z else None
        lib().orc_init_z(C.byref(m), bp(self.geno), ip(Z) if want_z else None, ip(nA), ip(nQ))
        return (nA, nQ, Z) if want_z else (nA, nQ)

    def update_p(self, nA, eps, F, lam, rep):
        m = self.model
        P = np.zeros((m.L, m.K, m.A))
        lib().orc_update_p(C.byref(m), ip(self.nall), ip(nA), dp(eps), dp(F), dp(lam), rep, dp(P))
        return P

    def update_q_gibbs(self, nQ, alpha, rep):
        m = self.model
        Q = np.zeros((m.N, m.K))
        lib().orc_update_q_gibbs(C.byref(m), ip(nQ), dp(alpha), rep, dp(Q))
        return Q

    def update_q_metro(self, P, alpha, rep, Q):
        m = self.model
        Q = np.ascontiguousarray(Q, dtype=np.float64).copy()
        acc = np.zeros(m.N, np.int32)
        lib().orc_update_q_metro(C.byref(m), bp(self.geno), dp(P), dp(alpha), rep, dp(Q), ip(acc))
        return Q, acc

    def calc_like(self, Q, P):
        return lib().orc_calc_like(C.byref(self.model), bp(self.geno), dp(Q), dp(P))

    def update_alpha(self, Q, alpha, rep):
        a = np.ascontiguousarray(alpha, dtype=np.float64).copy()
        lib().orc_update_alpha(C.byref(self.model), dp(Q), rep, dp(a))
        return a

    def update_lambda(self, P, eps, lam, rep):
        x = np.ascontiguousarray(lam, dtype=np.float64).copy()
        lib().orc_update_lambda(C.byref(self.model), ip(self.nall), dp(P), dp(eps), rep, dp(x))
        return x

    def update_eps(self, P, F, lam0, eps, rep):
        e = np.ascontiguousarray(eps, dtype=np.float64).copy()
        lib().orc_update_eps(C.byref(self.model), ip(self.nall), dp(P), dp(F), float(lam0), rep,
                             dp(e))
        return e

    def update_fst(self, eps, P, F, rep):
        f = np.ascontiguousarray(F, dtype=np.float64).copy()
        lib().orc_update_fst(C.byref(self.model), ip(self.nall), dp(eps), dp(P), rep, dp(f))
        return f
