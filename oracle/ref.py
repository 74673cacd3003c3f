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


# ---- the reference's parameter, data-file and result-file code -------------
# (params.c, datain.c, output.c; used to pin structure_amd/host/*.c)

# shost_params field -> reference global (structure.h:70-131)
PARAM_GLOBALS = {
    "numinds": "NUMINDS", "numloci": "NUMLOCI", "missing": "MISSING", "label": "LABEL",
    "popdata": "POPDATA", "popflag": "POPFLAG", "phenotype": "PHENOTYPE",
    "extracols": "EXTRACOLS", "markernames": "MARKERNAMES", "phaseinfo": "PHASEINFO",
    "mapdistances": "MAPDISTANCES", "markovphase": "MARKOVPHASE", "ploidy": "LINES",
    "onerowperind": "ONEROWPERIND", "recessivealleles": "RECESSIVEALLELES",
    "notambiguous": "NOTAMBIGUOUS", "phased": "PHASED", "locdata": "LOCDATA",
    "maxpops": "MAXPOPS", "burnin": "BURNIN", "numreps": "NUMREPS",
    "usepopinfo": "USEPOPINFO", "pfrompopflagonly": "PFROMPOPFLAGONLY",
    "inferalpha": "INFERALPHA", "inferlambda": "INFERLAMBDA",
    "popspecificlambda": "POPSPECIFICLAMBDA", "popalphas": "POPALPHAS",
    "computeprob": "COMPUTEPROB", "noadmix": "NOADMIX", "admburnin": "ADMBURNIN",
    "printqhat": "PRINTQHAT", "printqsum": "PRINTQSUM", "updatefreq": "UPDATEFREQ",
    "printlikes": "PRINTLIKES", "intermedsave": "INTERMEDSAVE", "printkld": "PRINTKLD",
    "printnet": "PRINTNET", "printlambda": "PRINTLAMBDA", "echodata": "ECHODATA",
    "ancestdist": "ANCESTDIST", "numboxes": "NUMBOXES", "gensback": "GENSBACK",
    "freqscorr": "FREQSCORR", "onefst": "ONEFST", "unifprioralpha": "UNIFPRIORALPHA",
    "startatpopinfo": "STARTATPOPINFO", "randomize": "RANDOMIZE", "linkage": "LINKAGE",
    "metrofreq": "METROFREQ", "reporthitrate": "REPORTHITRATE", "noqs": "NOQS",
    "posterior": "POSTERIOR", "picturefile": "PICTUREFILE", "individualr": "INDIVIDUALR",
    "numsimstats": "NUMSIMSTATS", "poolfreq": "POOLFREQ", "locusxonly": "LOCUSxONLY",
    "phenotypecol": "PHENOTYPECOL", "missingpheno": "MISSINGPHENO", "locprior": "LOCPRIOR",
    "updatelocprior": "UPDATELOCPRIOR", "locispop": "LOCISPOP",
    "printlocprior": "PRINTLOCPRIOR", "seed": "SEED", "origmissing": "ORIGMISSING",
}
PARAM_DOUBLES = {
    "ancestpint": "ANCESTPINT", "migrprior": "MIGRPRIOR", "alpha": "ALPHA",
    "log10rpropsd": "LOG10RPROPSD", "log10rmin": "LOG10RMIN", "log10rmax": "LOG10RMAX",
    "log10rstart": "LOG10RSTART", "fpriormean": "FPRIORMEAN", "fpriorsd": "FPRIORSD",
    "lambda_": "LAMBDA", "alphamax": "ALPHAMAX", "alphapriora": "ALPHAPRIORA",
    "alphapriorb": "ALPHAPRIORB", "alphapropsd": "ALPHAPROPSD", "emerror": "EMERROR",
    "locpriorinit": "LOCPRIORINIT", "maxlocprior": "MAXLOCPRIOR",
    "locpriorstep": "LOCPRIORSTEP",
}


def _str_global(name):
    return (C.c_char * 201).in_dll(lib(), name)


def get_params_globals() -> dict:
    out = {k: get_int(g) for k, g in PARAM_GLOBALS.items()}
    out.update({k: C.c_double.in_dll(lib(), g).value for k, g in PARAM_DOUBLES.items()})
    out["datafile"] = _str_global("DATAFILE").value.decode()
    out["outfile"] = _str_global("OUTFILE").value.decode()
    return out


def set_params_globals(d: dict) -> None:
    for k, g in PARAM_GLOBALS.items():
        if k in d:
            set_int(g, d[k])
    for k, g in PARAM_DOUBLES.items():
        if k in d:
            set_double(g, d[k])
    for k, g in (("datafile", "DATAFILE"), ("outfile", "OUTFILE")):
        if k in d:
            _str_global(g).value = d[k].encode()
    set_int("NOALPHA", 0)


def _argv(args):
    return len(args), (C.c_char_p * (len(args) + 1))(*[a.encode() for a in args], None)


def get_params(args) -> dict:
    """GetParams(0, argc, argv) (params.c:49-82); the reference exits the
    process on a missing file or key, so callers pass valid inputs only."""
    L = lib()
    L.GetParams.restype = None
    L.GetParams.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_char_p)]
    n, arr = _argv(list(args))
    L.GetParams(0, n, arr)
    return get_params_globals()


def read_input(params: dict):
    """ReadInputFile + FindMaxAlleles + CountAlleles (datain.c) on the file and
    options in params; returns (geno [N][2][L], nall, translation [L][A],
    labels, population, popflag, maxalleles, missing after recoding)."""
    L = lib()
    set_params_globals(params)
    N, Lc = params["numinds"], params["numloci"]
    geno = np.zeros((N, 2, Lc), np.int32)
    mapd = np.zeros(Lc)
    marker = (C.c_char * (15 * Lc))()
    inds = individuals(N)
    phase = np.zeros(N * Lc)
    rec = np.zeros(Lc, np.int32)
    L.ReadInputFile.restype = None
    L.ReadInputFile.argtypes = [_I, _D, C.c_void_p, C.c_void_p, _D, _I]
    L.ReadInputFile(ip(geno), dp(mapd), marker, inds, dp(phase), ip(rec))
    L.FindMaxAlleles.restype = C.c_int
    L.FindMaxAlleles.argtypes = [_I, _I]
    A = L.FindMaxAlleles(ip(geno), ip(rec))
    set_int("MAXALLELES", A)
    nall = np.zeros(Lc, np.int32)
    trans = np.zeros((Lc, max(A, 1)), np.int32)
    L.CountAlleles(ip(geno), ip(nall), ip(trans), ip(rec))
    raw = bytes(inds)
    labels, pops, flags = [], [], []
    for i in range(N):
        rec_i = raw[i * IND_SIZE:(i + 1) * IND_SIZE]
        labels.append(rec_i[:12].split(b"\0")[0].decode())
        pops.append(int.from_bytes(rec_i[12:16], "little", signed=True))
        flags.append(int.from_bytes(rec_i[16:20], "little", signed=True))
    names = [bytes(marker)[j * 15:(j + 1) * 15].split(b"\0")[0].decode() for j in range(Lc)]
    return dict(geno=geno, nall=nall, translation=trans, labels=labels,
                population=np.array(pops, np.int32), popflag=np.array(flags, np.int32),
                maxalleles=A, missing=get_int("MISSING"), markernames=names,
                origmissing=get_int("ORIGMISSING"))


def output_results(params: dict, inp: dict, rep, qsum, psum, epssum, fstsum, alphasum,
                   lambdasum, sumlikes, sumsqlikes, args=("structure",), ancestdist=None,
                   final=1, savefreq=0) -> None:
    """OutPutResults (output.c:1354-1423) for the state given; final=1 writes
    <OUTFILE>_f (and _q with PRINTQHAT), final=0 the intermediate file
    <OUTFILE>_<(rep - BURNIN) / savefreq>."""
    L = lib()
    set_params_globals(params)
    set_int("MAXALLELES", inp["maxalleles"])
    set_int("MISSING", inp["missing"])
    set_int("ORIGMISSING", inp["origmissing"])
    N = params["numinds"]
    inds = individuals(N)
    buf = bytearray(IND_SIZE * N)
    for i in range(N):
        lab = inp["labels"][i].encode()[:11]
        buf[i * IND_SIZE:i * IND_SIZE + len(lab)] = lab
        buf[i * IND_SIZE + 12:i * IND_SIZE + 16] = int(inp["population"][i]).to_bytes(
            4, "little", signed=True)
        buf[i * IND_SIZE + 16:i * IND_SIZE + 20] = int(inp["popflag"][i]).to_bytes(
            4, "little", signed=True)
    C.memmove(inds, bytes(buf), len(buf))
    marker = (C.c_char * (15 * len(inp["markernames"])))()
    mb = bytearray(15 * len(inp["markernames"]))
    for j, nm in enumerate(inp["markernames"]):
        e = nm.encode()[:14]
        mb[j * 15:j * 15 + len(e)] = e
    C.memmove(marker, bytes(mb), len(mb))
    arrs = [np.ascontiguousarray(a, dtype=np.float64)
            for a in (psum, qsum, fstsum, alphasum, epssum, lambdasum)]
    geno = np.ascontiguousarray(inp["geno"], dtype=np.int32)
    nall = np.ascontiguousarray(inp["nall"], dtype=np.int32)
    trans = np.ascontiguousarray(inp["translation"], dtype=np.int32)
    zeros = np.zeros(max(N, 8))
    vp = C.c_void_p
    L.OutPutResults.restype = None
    L.OutPutResults.argtypes = [_I, C.c_int, C.c_int, vp, _D, _D, _D, vp, vp, C.c_double,
                                C.c_double, _D, _D, _D, _I, _I, C.c_int, vp, _D, _D, _D, _D,
                                _D, C.c_int, _D, _D, C.c_int, C.POINTER(C.c_char_p)]
    lam = np.ascontiguousarray(lambdasum, dtype=np.float64)
    n, argv = _argv(list(args))
    psum_, qsum_, fst_, alpha_, eps_, lsum_ = arrs
    ad = None if ancestdist is None else np.ascontiguousarray(ancestdist, dtype=np.int32)
    L.OutPutResults(ip(geno), rep, savefreq, inds, dp(psum_), dp(qsum_), dp(fst_),
                    None if ad is None else ip(ad), None,
                    sumlikes, sumsqlikes, dp(alpha_), dp(zeros), dp(zeros), ip(nall), ip(trans),
                    final, marker, dp(zeros), dp(eps_), dp(lam), dp(lsum_), dp(zeros), 0,
                    dp(zeros), dp(zeros), n, argv)
