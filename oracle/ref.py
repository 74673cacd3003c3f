"""ctypes front end of oracle/_ref/libstructref.so -- test infrastructure only.

``libstructref.so`` is the reference's own CPU code (pmelsted/structure
``src/*.c`` minus ``structure.c``'s ``main``) compiled from
``/root/reference/src`` by ``oracle/Makefile``, with ``rand()`` redirected to
the queue in ``oracle/ref_harness.c``.  Tests use it to pin the oracle's
restatement; it is never part of the product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(_HERE, "_ref", "libstructref.so")
_LIB = None

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
IND_SIZE = 32   # sizeof(struct IND), structure.h:52-59: char[12] + 5 ints


def available() -> bool:
    return os.path.exists(PATH)


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = C.CDLL(PATH)
        L = _LIB
        vp = C.c_void_p
        sig = {
            "UpdateZ": (None, [_I, _D, _D, _I, _D]),
            "GetNumFromPops": (None, [_I, _I, _I, _I, vp]),
            "GetNumLociPop": (None, [_I, _I, C.c_int]),
            "CalcLike": (C.c_double, [_I, _I, _D, _D, _I, _D, _D]),
            "CalcLikeIndDiff": (C.c_double, [_I, _I, _D, _D, _D, C.c_int, _I]),
            "LogProbQ": (C.c_double, [_D, C.c_double, vp]),
            "LogProbQonepop": (C.c_double, [_D, C.c_double, C.c_double, vp, C.c_int]),
            "AlphaPriorDiff": (C.c_double, [C.c_double, C.c_double]),
            "UpdateAlpha": (None, [_D, _D, vp, C.c_int]),
            "UpdateLambda": (None, [_D, _D, _D, _I]),
            "UpdatePopLambda": (None, [_D, _D, _I]),
            "NonIndependenceUpdateEpsilon": (None, [_D, _D, _D, _I, C.c_double]),
            "FPriorDiff": (C.c_double, [C.c_double, C.c_double]),
            "FlikeFreqs": (C.c_double, [C.c_double, _D, _D, _I, C.c_int]),
            "UpdateFst": (None, [_D, _D, _D, _I]),
            "mylgamma": (C.c_double, [C.c_double]),
            "snorm": (C.c_double, []),
            "rnd": (C.c_double, []),
            "UpdateSums": (None, [_D, _D, _I, _D, _D, _D, _D, _I, _I, _D, _D, _D, _D, _D, _D,
                                  C.c_int]),
            "InitFreqPriors": (None, [_D, _D, _I, _I]),
            "InitializeZ": (None, [_I, vp, _I]),
            "CountAlleles": (None, [_I, _I, _I, _I]),
            "UpdateP": (None, [_D, _D, _D, _I, _I, _I, _D, vp, _D]),
            "UpdateQAdmixture": (None, [_D, _I, _D, vp, _D]),
            "UpdateQMetro": (None, [_I, _I, _D, _D, _D, C.c_int, vp, _I, _D]),
            "SampleVar": (C.c_double, [C.c_double, C.c_double, C.c_long]),
            "EstLogProb": (C.c_double, [C.c_double, C.c_double, C.c_int]),
            "sref_set_rand": (None, [_I, C.c_size_t]),
            "sref_rand_consumed": (C.c_size_t, []),
            "sref_rand_overrun": (C.c_size_t, []),
            "sref_rand_max": (C.c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
    return _LIB


def set_int(name: str, v: int) -> None:
    C.c_int.in_dll(lib(), name).value = int(v)


def set_double(name: str, v: float) -> None:
    C.c_double.in_dll(lib(), name).value = float(v)


def get_int(name: str) -> int:
    return C.c_int.in_dll(lib(), name).value


def configure(N, L, K, A, **kw) -> None:
    """Set the reference globals (structure.h:70-131) for an admixture run."""
    base = dict(NUMINDS=N, NUMLOCI=L, MAXPOPS=K, MAXALLELES=A, LINES=2, MISSING=-9,
                RECESSIVEALLELES=0, USEPOPINFO=0, LOCPRIOR=0, NOADMIX=0, LINKAGE=0,
                PFROMPOPFLAGONLY=0, POPALPHAS=0, UNIFPRIORALPHA=1, FREQSCORR=1, ONEFST=0,
                ANCESTDIST=0, NUMBOXES=1000, STARTATPOPINFO=0, POPDATA=0, ADMBURNIN=500,
                BURNIN=0, INFERLAMBDA=0, POPSPECIFICLAMBDA=0, REPORTHITRATE=0, METROFREQ=10,
                UPDATEFREQ=100)
    dbl = dict(ALPHAPROPSD=0.025, ALPHAMAX=10.0, ALPHAPRIORA=1.0, ALPHAPRIORB=2.0,
               FPRIORMEAN=0.01, FPRIORSD=0.05, LAMBDA=1.0, ALPHA=1.0)
    for k, v in kw.items():
        if k in dbl:
            dbl[k] = v
        else:
            base[k] = v
    for k, v in base.items():
        set_int(k, v)
    for k, v in dbl.items():
        set_double(k, v)


def set_rand(vals) -> None:
    a = np.ascontiguousarray(vals, dtype=np.int32)
    lib().sref_set_rand(a.ctypes.data_as(_I), a.size)


def individuals(N: int) -> C.Array:
    return (C.c_char * (IND_SIZE * N))()


def dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def ip(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_I)
