"""ctypes binding of the host side of the `structure` command
(include/structure_host.h, include/structure_align.h, in libsgpu.so): the
mainparams/extraparams reader, the genotype file reader with allele
recoding, the result-file writer, and replicate label alignment."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import SgpuError, lib

STRLEN, GENELEN, LABELLEN = 200, 15, 12

_I = C.c_int32
_D = C.c_double

# field order of shost_params (include/structure_host.h)
_PARAM_FIELDS = (
    [("datafile", C.c_char * (STRLEN + 1)), ("outfile", C.c_char * (STRLEN + 1))]
    + [(n, _I) for n in ("numinds", "numloci", "missing", "label", "popdata", "popflag",
                         "phenotype", "extracols", "markernames", "phaseinfo", "mapdistances",
                         "markovphase", "ploidy", "onerowperind", "recessivealleles",
                         "notambiguous", "phased", "locdata", "maxpops", "burnin", "numreps",
                         "usepopinfo", "pfrompopflagonly", "inferalpha", "inferlambda",
                         "popspecificlambda", "popalphas", "computeprob", "noadmix",
                         "admburnin", "printqhat", "printqsum", "updatefreq", "printlikes",
                         "intermedsave", "printkld", "printnet", "printlambda", "echodata",
                         "ancestdist", "numboxes", "gensback", "freqscorr", "onefst",
                         "unifprioralpha", "startatpopinfo", "randomize", "linkage",
                         "metrofreq", "reporthitrate")]
    + [(n, _D) for n in ("ancestpint", "migrprior", "alpha", "log10rpropsd", "log10rmin",
                         "log10rmax", "log10rstart", "fpriormean", "fpriorsd", "lambda_",
                         "alphamax", "alphapriora", "alphapriorb", "alphapropsd")]
    + [(n, _I) for n in ("noqs", "posterior", "picturefile", "individualr", "numsimstats",
                         "poolfreq", "locusxonly", "phenotypecol", "missingpheno")]
    + [("emerror", _D)]
    + [(n, _I) for n in ("locprior", "updatelocprior", "locispop", "printlocprior")]
    + [(n, _D) for n in ("locpriorinit", "maxlocprior", "locpriorstep")]
    + [("seed", _I), ("origmissing", _I)])


class Params(C.Structure):
    """Mirror of ``shost_params``."""
    _fields_ = _PARAM_FIELDS

    def as_dict(self):
        out = {}
        for n, t in self._fields_:
            v = getattr(self, n)
            out[n] = v.decode() if isinstance(v, bytes) else v
        return out


class Input(C.Structure):
    """Mirror of ``shost_input``."""
    _fields_ = [("N", _I), ("L", _I), ("lines", _I), ("maxalleles", _I), ("missing", _I),
                ("geno", C.POINTER(_I)), ("nall", C.POINTER(_I)),
                ("translation", C.POINTER(_I)), ("label", C.c_void_p),
                ("population", C.POINTER(_I)), ("popflag", C.POINTER(_I)),
                ("phenotype", C.POINTER(_I)), ("location", C.POINTER(_I)),
                ("markername", C.c_void_p), ("mapdistance", C.POINTER(_D))]


class Results(C.Structure):
    """Mirror of ``shost_results``."""
    _fields_ = [("rep", _I), ("qsum", C.POINTER(_D)), ("psum", C.POINTER(_D)),
                ("epssum", C.POINTER(_D)), ("fstsum", C.POINTER(_D)),
                ("alphasum", C.POINTER(_D)), ("lambdasum", C.POINTER(_D)),
                ("sumlikes", _D), ("sumsqlikes", _D), ("ancestdist", C.POINTER(_I)),
                ("saveindex", _I)]


_SIGS = {
    "shost_preset": (None, [C.POINTER(Params)]),
    "shost_read_params": (C.c_int, [C.POINTER(Params), C.c_char_p]),
    "shost_command_line": (C.c_int, [C.POINTER(Params), C.c_int, C.POINTER(C.c_char_p)]),
    "shost_get_params": (C.c_int, [C.POINTER(Params), C.c_int, C.POINTER(C.c_char_p)]),
    "shost_param_file_name": (None, [C.c_int, C.POINTER(C.c_char_p), C.c_char_p, C.c_char_p,
                                     C.c_size_t]),
    "shost_check": (C.c_int, [C.POINTER(Params)]),
    "shost_update_freq": (_I, [C.POINTER(Params)]),
    "shost_read_input": (C.c_int, [C.POINTER(Params), C.POINTER(Input)]),
    "shost_free_input": (None, [C.POINTER(Input)]),
    "shost_geno_int8": (C.c_int, [C.POINTER(Input), C.c_void_p]),
    "shost_write_results": (C.c_int, [C.POINTER(Params), C.POINTER(Input), C.POINTER(Results),
                                      C.c_int, C.c_int, C.POINTER(C.c_char_p)]),
    "shost_est_log_prob": (_D, [_D, _D, C.c_int]),
    "shost_save_freq": (_I, [C.POINTER(Params)]),
    "shost_calc_like": (_D, [C.POINTER(Input), C.c_int, C.c_void_p, C.c_void_p]),
    "shost_assign": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "shost_align_q": (C.c_int, [C.c_void_p, C.c_void_p, _I, _I, C.c_void_p]),
    "shost_permute_clusters": (C.c_int, [C.c_void_p, _I, _I, _I, C.c_void_p, C.c_void_p]),
    "shost_replicate_summary": (C.c_int, [C.c_void_p, C.c_void_p, _I, _I, _I, C.c_void_p,
                                          C.c_void_p, C.POINTER(_D), C.POINTER(_D),
                                          C.c_void_p]),
}
_READY = False


def _lib():
    global _READY
    L = lib()
    if not _READY:
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _READY = True
    return L


def _argv(args):
    arr = (C.c_char_p * (len(args) + 1))(*[a.encode() for a in args], None)
    return len(args), arr


def _vp(a):
    return a.ctypes.data_as(C.c_void_p)


def preset() -> Params:
    p = Params()
    _lib().shost_preset(C.byref(p))
    return p


def read_params(path, p: Params | None = None) -> Params:
    p = p if p is not None else preset()
    rc = _lib().shost_read_params(C.byref(p), str(path).encode())
    if rc:
        raise SgpuError(rc, f"cannot read {path}")
    return p


def get_params(args) -> Params:
    """GetParams: preset, -m / -e files (defaults mainparams / extraparams in
    the working directory), command line, checks; args includes argv[0]."""
    p = Params()
    n, arr = _argv(list(args))
    rc = _lib().shost_get_params(C.byref(p), n, arr)
    if rc:
        raise SgpuError(rc, "parameter error (see stdout)")
    return p


def param_file_name(args, flag) -> str:
    n, arr = _argv(list(args))
    out = C.create_string_buffer(500)
    _lib().shost_param_file_name(n, arr, flag.encode(), out, 500)
    return out.value.decode()


class InputFile:
    """A genotype file read and recoded (``shost_read_input``)."""

    def __init__(self, p: Params):
        self._in = Input()
        self.params = p
        rc = _lib().shost_read_input(C.byref(p), C.byref(self._in))
        if rc:
            raise SgpuError(rc, f"cannot read {p.datafile.decode()}")
        s = self._in
        N, L, lines, A = s.N, s.L, s.lines, s.maxalleles
        self.N, self.L, self.A, self.missing = N, L, A, s.missing
        self.geno = np.ctypeslib.as_array(s.geno, (N, lines, L)).copy()
        self.nall = np.ctypeslib.as_array(s.nall, (L,)).copy()
        self.translation = (np.ctypeslib.as_array(s.translation, (L, max(A, 1))).copy()
                            if A > 0 else np.zeros((L, 0), np.int32))
        raw = C.string_at(s.label, N * LABELLEN)
        self.labels = [raw[i * LABELLEN:(i + 1) * LABELLEN].split(b"\0")[0].decode()
                       for i in range(N)]
        self.population = np.ctypeslib.as_array(s.population, (N,)).copy()
        self.popflag = np.ctypeslib.as_array(s.popflag, (N,)).copy()
        self.phenotype = np.ctypeslib.as_array(s.phenotype, (N,)).copy()
        raw = C.string_at(s.markername, L * GENELEN)
        self.markernames = [raw[j * GENELEN:(j + 1) * GENELEN].split(b"\0")[0].decode()
                            for j in range(L)]

    def geno_int8(self) -> np.ndarray:
        out = np.empty((self.N, 2, self.L), np.int8)
        rc = _lib().shost_geno_int8(C.byref(self._in), _vp(out))
        if rc:
            raise SgpuError(rc, "alleles do not fit the packed formats")
        return out

    def geno16(self) -> np.ndarray:
        """Recoded genotypes [N][2][L] as int16 with -9 for missing (any allele count)."""
        g = self.geno.astype(np.int16)
        g[self.geno == self.missing] = -9
        return g

    def write_results(self, rep, qsum, psum, epssum, fstsum, alphasum, lambdasum, sumlikes,
                      sumsqlikes, final=1, args=("structure",), params: Params | None = None,
                      ancestdist=None, saveindex=0):
        p = params if params is not None else self.params
        keep = [np.ascontiguousarray(a, dtype=np.float64)
                for a in (qsum, psum, epssum, fstsum, alphasum, lambdasum)]
        ptr = [a.ctypes.data_as(C.POINTER(_D)) for a in keep]
        ad = None if ancestdist is None else np.ascontiguousarray(ancestdist, dtype=np.int32)
        r = Results(rep, *ptr, sumlikes, sumsqlikes,
                    None if ad is None else ad.ctypes.data_as(C.POINTER(_I)), saveindex)
        n, arr = _argv(list(args))
        rc = _lib().shost_write_results(C.byref(p), C.byref(self._in), C.byref(r), final, n, arr)
        if rc:
            raise SgpuError(rc, "cannot write results")

    def calc_like(self, Q, P) -> float:
        """CalcLike (CalcLike.c:387-426) on the host, as PRINTLIKES uses it."""
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        P = np.ascontiguousarray(P, dtype=np.float64)
        return _lib().shost_calc_like(C.byref(self._in), Q.shape[1], _vp(Q), _vp(P))

    def close(self):
        if self._in is not None:
            _lib().shost_free_input(C.byref(self._in))
            self._in = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def est_log_prob(sumlikes, sumsqlikes, reps) -> float:
    return _lib().shost_est_log_prob(sumlikes, sumsqlikes, reps)


def save_freq(p: Params) -> int:
    """INTERMEDSAVE's interval between intermediate result files (Init.c:377-385)."""
    return _lib().shost_save_freq(C.byref(p))


def assign(cost: np.ndarray) -> np.ndarray:
    c = np.ascontiguousarray(cost, dtype=np.float64)
    n = c.shape[0]
    out = np.empty(n, np.int32)
    rc = _lib().shost_assign(_vp(c), n, _vp(out))
    if rc:
        raise SgpuError(rc, "assignment failed")
    return out


def align_q(qref: np.ndarray, q: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(qref, dtype=np.float64)
    b = np.ascontiguousarray(q, dtype=np.float64)
    N, K = a.shape
    perm = np.empty(K, np.int32)
    rc = _lib().shost_align_q(_vp(a), _vp(b), N, K, _vp(perm))
    if rc:
        raise SgpuError(rc, "alignment failed")
    return perm


def replicate_summary(qbars: np.ndarray, lnp: np.ndarray):
    """Align R replicates [R][N][K] to the first; returns (qmean, qsd, lnp_mean,
    lnp_se, perms)."""
    q = np.ascontiguousarray(qbars, dtype=np.float64)
    lp = np.ascontiguousarray(lnp, dtype=np.float64)
    R, N, K = q.shape
    qmean = np.empty((N, K))
    qsd = np.empty((N, K))
    perms = np.empty((R, K), np.int32)
    m, se = _D(), _D()
    rc = _lib().shost_replicate_summary(_vp(q), _vp(lp), R, N, K, _vp(qmean), _vp(qsd),
                                        C.byref(m), C.byref(se), _vp(perms))
    if rc:
        raise SgpuError(rc, "replicate summary failed")
    return qmean, qsd, m.value, se.value, perms
