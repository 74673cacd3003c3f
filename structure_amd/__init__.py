"""structure_amd: MI355X (gfx950) admixture-MCMC sweep of STRUCTURE.

Python mirror of the reference's operator interface for the hot path: the
``*CL`` update wrappers (UpdatePCL, UpdateQCL, UpdateZCL, UpdateAlphaCL,
UpdateEpsilonCL, UpdateFstCL, DataCollectionCL) and the buffer read/write of
``Kernels.c``, on top of the C-ABI in ``include/sgpu.h`` implemented by
``libsgpu.so`` (hand-written HIP kernels).  There is no CPU fallback: if the
library or a GPU is missing, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsgpu.so")
_LIB = None

OK, ERR_ARG, ERR_NOMEM, ERR_DEVICE, ERR_STATE, ERR_UNSUPPORTED = 0, -1, -2, -3, -4, -5
GENO_SNP2, GENO_BYTE = 1, 2
(Q, P, EPS, FST, ALPHA, LAMBDA, NA, NQ, QSUM, PSUM, FSTSUM, EPSSUM, ALPHASUM, LAMBDASUM, LIKE,
 NSUM, ANCESTDIST) = range(1, 18)
Z_LIKE = 1
PHASES = ("P", "Q", "Z", "alpha", "lambda", "eps", "fst", "collect")


class SgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sgpu error {code}: {msg}")
        self.code = code


class Model(C.Structure):
    """Mirror of ``sgpu_model`` (include/sgpu.h)."""

    _fields_ = [(n, C.c_int32) for n in (
        "N", "L", "K", "A", "burnin", "numreps", "freqscorr", "onefst", "inferalpha",
        "popalphas", "unifprioralpha", "inferlambda", "popspecificlambda", "metrofreq",
        "computeprob")] + [(n, C.c_double) for n in (
            "alpha0", "lambda0", "alphapropsd", "alphamax", "alphapriora", "alphapriorb",
            "fpriormean", "fpriorsd")] + [("seed", C.c_uint32), ("chain", C.c_uint32)] + [
        (n, C.c_int32) for n in ("ancestdist", "numboxes", "pfrompopflagonly", "reserved")]


class Data(C.Structure):
    """Mirror of ``sgpu_data``; popflag may be null."""

    _fields_ = [("format", C.c_int32), ("geno", C.POINTER(C.c_uint32)),
                ("geno_words", C.c_size_t), ("nall", C.POINTER(C.c_int32)),
                ("popflag", C.POINTER(C.c_int32))]


class Tuning(C.Structure):
    """Mirror of ``sgpu_tuning``: block targets of the batched UpdateZ and
    Metropolis launches (0 = library default)."""

    _fields_ = [("z_blocks", C.c_int32), ("metro_blocks", C.c_int32),
                ("byte_loci", C.c_int32), ("byte_chunk", C.c_int32)]


# Defaults of the shipped extraparams / PresetValues (extraparams:1-120, params.c:571-670)
DEFAULTS = dict(burnin=0, numreps=0, freqscorr=1, onefst=0, inferalpha=1, popalphas=0,
                unifprioralpha=1, inferlambda=0, popspecificlambda=0, metrofreq=10,
                computeprob=1, alpha0=1.0, lambda0=1.0, alphapropsd=0.025, alphamax=10.0,
                alphapriora=1.0, alphapriorb=2.0, fpriormean=0.01, fpriorsd=0.05, seed=1,
                chain=0, ancestdist=0, numboxes=1000, pfrompopflagonly=0, reserved=0)


def make_model(N, L, K, A, **kw) -> Model:
    d = dict(DEFAULTS)
    d.update(kw)
    return Model(N=N, L=L, K=K, A=A, **d)


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise SgpuError(ERR_STATE, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        u32, i32, sz = C.c_uint32, C.c_int32, C.c_size_t
        sig = {
            "sgpu_packed_words": (sz, [i32, i32, i32]),
            "sgpu_pack_snp2": (C.c_int, [vp, i32, i32, vp, sz]),
            "sgpu_pack_byte": (C.c_int, [vp, i32, i32, vp, sz]),
            "sgpu_pack_byte32": (C.c_int, [vp, i32, i32, i32, vp, sz]),
            "sgpu_create": (C.c_int, [C.POINTER(Model), C.POINTER(Data), C.POINTER(vp)]),
            "sgpu_create_batch": (C.c_int, [C.POINTER(Model), i32, C.POINTER(Data), i32, vp,
                                            C.POINTER(Tuning), C.POINTER(vp)]),
            "sgpu_num_chains": (C.c_int, [vp]),
            "sgpu_get_chain_state": (C.c_int, [vp, i32, i32, vp, sz]),
            "sgpu_set_chain_state": (C.c_int, [vp, i32, i32, vp, sz]),
            "sgpu_destroy": (C.c_int, [vp]),
            "sgpu_init_state": (C.c_int, [vp]),
            "sgpu_get_state": (C.c_int, [vp, i32, vp, sz]),
            "sgpu_set_state": (C.c_int, [vp, i32, vp, sz]),
            "sgpu_update_p": (C.c_int, [vp, u32, i32]),
            "sgpu_update_q": (C.c_int, [vp, u32, i32]),
            "sgpu_update_z": (C.c_int, [vp, u32, i32, vp]),
            "sgpu_update_alpha": (C.c_int, [vp, u32]),
            "sgpu_update_lambda": (C.c_int, [vp, u32]),
            "sgpu_update_eps": (C.c_int, [vp, u32, i32]),
            "sgpu_update_fst": (C.c_int, [vp, u32]),
            "sgpu_collect": (C.c_int, [vp, u32]),
            "sgpu_sweep": (C.c_int, [vp, u32]),
            "sgpu_run_sweeps": (C.c_int, [vp, u32, u32]),
            "sgpu_synchronize": (C.c_int, [vp]),
            "sgpu_set_timing": (C.c_int, [vp, i32]),
            "sgpu_get_timing": (C.c_int, [vp, vp, vp]),
            "sgpu_eval_math": (C.c_int, [i32, vp, vp, C.c_int64]),
            "sgpu_eval_math_noalloc": (C.c_int, [i32, vp, vp, C.c_int64]),
            "sgpu_last_error": (C.c_char_p, []),
            "sgpu_version": (C.c_char_p, []),
            "sgpu_synth_snp": (C.c_int, [i32, i32, i32, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_uint64, i32, vp, sz, vp, vp, vp]),
            "sgpu_synth_microsat": (C.c_int, [i32, i32, i32, i32, i32, C.c_double, C.c_double,
                                              C.c_uint64, vp, vp, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(rc):
    if rc != OK:
        raise SgpuError(rc, lib().sgpu_last_error().decode())
    return rc


def _vp(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


FN_LOG, FN_COS, FN_LGAMMA, FN_SIN = 0, 1, 2, 3


EVAL_MATH_NOALLOC_MAX = 65536   # SGPU_EVAL_MATH_NOALLOC_MAX


def eval_math(fn: int, x, noalloc: bool = False) -> np.ndarray:
    """The device's own log / cos / sin / lgamma over an array (diagnostics);
    noalloc: through the allocation-free entry point (at most
    EVAL_MATH_NOALLOC_MAX values)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    f = lib().sgpu_eval_math_noalloc if noalloc else lib().sgpu_eval_math
    _check(f(fn, _vp(x), _vp(y), x.size))
    return y


# ---- genotype packing and synthetic data -------------------------------------

def packed_words(fmt, N, L) -> int:
    return lib().sgpu_packed_words(fmt, N, L)


def pack(geno: np.ndarray, fmt: int | None = None):
    """Pack recoded genotypes [N][2][L] (-9 missing; alleles up to 254 in any
    integer type).  fmt=None picks SNP2 when every site is representable, else
    the byte format.  Returns (fmt, words)."""
    geno = np.asarray(geno)
    N, two, L = geno.shape
    assert two == 2
    wide = geno.dtype != np.int8 and geno.size > 0 and int(geno.max()) > 126
    g8 = None if wide else np.ascontiguousarray(geno, dtype=np.int8)
    order = [GENO_SNP2, GENO_BYTE] if fmt is None else [fmt]
    if wide:
        order = [f for f in order if f == GENO_BYTE] or order
    for f in order:
        out = np.empty(packed_words(f, N, L), np.uint32)
        if wide and f == GENO_BYTE:
            g32 = np.ascontiguousarray(geno, dtype=np.int32)
            rc = lib().sgpu_pack_byte32(_vp(g32), N, L, -9, _vp(out), out.size)
        elif wide:
            rc = ERR_UNSUPPORTED   # alleles above 1 are not 2-bit data
        else:
            fn = lib().sgpu_pack_snp2 if f == GENO_SNP2 else lib().sgpu_pack_byte
            rc = fn(_vp(g8), N, L, _vp(out), out.size)
        if rc == OK:
            return f, out
        if fmt is not None or rc != ERR_UNSUPPORTED:
            _check(rc)
    raise SgpuError(ERR_UNSUPPORTED, "genotypes not packable")


def synth_snp(N, L, Ktrue=5, flo=0.01, fhi=0.1, qalpha=0.3, miss=0.01, seed=1, threads=0,
              want_geno=False, want_q=False):
    words = np.empty(packed_words(GENO_SNP2, N, L), np.uint32)
    nall = np.empty(L, np.int32)
    geno = np.empty((N, 2, L), np.int8) if want_geno else None
    qt = np.empty((N, Ktrue)) if want_q else None
    _check(lib().sgpu_synth_snp(N, L, Ktrue, flo, fhi, qalpha, miss, seed, threads, _vp(words),
                                words.size, _vp(geno) if want_geno else None, _vp(nall),
                                _vp(qt) if want_q else None))
    return dict(words=words, nall=nall, geno=geno, qtrue=qt, A=2, fmt=GENO_SNP2)


def synth_microsat(N, L, amin=2, amax=10, Ktrue=3, qalpha=0.5, miss=0.02, seed=1, want_q=False):
    geno = np.empty((N, 2, L), np.int8)
    nall = np.empty(L, np.int32)
    qt = np.empty((N, Ktrue)) if want_q else None
    A = lib().sgpu_synth_microsat(N, L, amin, amax, Ktrue, qalpha, miss, seed, _vp(geno),
                                  _vp(nall), _vp(qt) if want_q else None)
    if A < 0:
        _check(A)
    return dict(geno=geno, nall=nall, qtrue=qt, A=A)


# ---- chain context ----------------------------------------------------------

def _state_shape(m: Model, which):
    N, L, K, A = m.N, m.L, m.K, m.A
    return {Q: ((N, K), np.float64), P: ((L, K, A), np.float64), EPS: ((L, A), np.float64),
            FST: ((K,), np.float64), ALPHA: ((K,), np.float64), LAMBDA: ((K,), np.float64),
            NA: ((L, K, A), np.int32), NQ: ((N, K), np.int32), QSUM: ((N, K), np.float64),
            PSUM: ((L, K, A), np.float64), FSTSUM: ((K,), np.float64),
            EPSSUM: ((L, A), np.float64), ALPHASUM: ((K,), np.float64),
            LAMBDASUM: ((K,), np.float64), LIKE: ((3,), np.float64),
            NSUM: ((1,), np.int64), ANCESTDIST: ((N, K, m.numboxes), np.int32)}[which]


class _Context:
    """Phase entry points shared by single chains and batches.  Each call
    advances every chain of the context."""

    _h = None
    models: list

    def close(self):
        if getattr(self, "_h", None):
            lib().sgpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # state of one chain
    def get_chain(self, chain, which) -> np.ndarray:
        shape, dt = _state_shape(self.models[chain], which)
        out = np.empty(shape, dt)
        _check(lib().sgpu_get_chain_state(self._h, chain, which, _vp(out), out.nbytes))
        return out

    def set_chain(self, chain, which, value):
        shape, dt = _state_shape(self.models[chain], which)
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=dt), shape))
        _check(lib().sgpu_set_chain_state(self._h, chain, which, _vp(a), a.nbytes))

    # phases
    def init(self):
        _check(lib().sgpu_init_state(self._h))

    def update_p(self, rep, collect=False):
        _check(lib().sgpu_update_p(self._h, rep, int(collect)))

    def update_q(self, rep, collect=False):
        _check(lib().sgpu_update_q(self._h, rep, int(collect)))

    def update_z(self, rep, like=False, want_z=False):
        m = self.models[0]
        z = np.empty((m.N, 2, m.L), np.int8) if want_z else None
        _check(lib().sgpu_update_z(self._h, rep, Z_LIKE if like else 0,
                                   _vp(z) if want_z else None))
        return z

    def update_alpha(self, rep):
        _check(lib().sgpu_update_alpha(self._h, rep))

    def update_lambda(self, rep):
        _check(lib().sgpu_update_lambda(self._h, rep))

    def update_eps(self, rep, collect=False):
        _check(lib().sgpu_update_eps(self._h, rep, int(collect)))

    def update_fst(self, rep):
        _check(lib().sgpu_update_fst(self._h, rep))

    def collect(self, rep):
        _check(lib().sgpu_collect(self._h, rep))

    def sweep(self, rep):
        _check(lib().sgpu_sweep(self._h, rep))

    def run(self, rep0, n):
        _check(lib().sgpu_run_sweeps(self._h, rep0, n))

    def synchronize(self):
        _check(lib().sgpu_synchronize(self._h))

    def set_timing(self, enable=True):
        _check(lib().sgpu_set_timing(self._h, int(enable)))

    def get_timing(self):
        """{phase: (total ms, launches)} since the last call (HIP events on the context stream)."""
        ms = np.zeros(len(PHASES))
        n = np.zeros(len(PHASES), np.int64)
        _check(lib().sgpu_get_timing(self._h, _vp(ms), _vp(n)))
        return {p: (float(ms[i]), int(n[i])) for i, p in enumerate(PHASES)}


class Chain(_Context):
    """One MCMC chain on the GPU (an ``sgpu_ctx``).

    Method names follow the reference's host wrappers: ``update_p`` ~ UpdatePCL,
    ``update_q`` ~ UpdateQCL, ``update_z`` ~ UpdateZCL, ``update_alpha`` ~
    UpdateAlphaCL, ``update_lambda`` ~ UpdateLambda, ``update_eps`` ~
    UpdateEpsilonCL, ``update_fst`` ~ UpdateFstCL, ``collect`` ~ DataCollectionCL;
    ``get``/``set`` ~ readBuffer/writeBuffer."""

    def __init__(self, model: Model, fmt: int, words: np.ndarray, nall: np.ndarray,
                 popflag: np.ndarray | None = None):
        self.model = model
        self.models = [model]
        self._words = np.ascontiguousarray(words, dtype=np.uint32)
        self._nall = np.ascontiguousarray(nall, dtype=np.int32)
        self._popflag = (None if popflag is None
                         else np.ascontiguousarray(popflag, dtype=np.int32))
        data = Data(format=fmt, geno=self._words.ctypes.data_as(C.POINTER(C.c_uint32)),
                    geno_words=self._words.size,
                    nall=self._nall.ctypes.data_as(C.POINTER(C.c_int32)),
                    popflag=(None if self._popflag is None
                             else self._popflag.ctypes.data_as(C.POINTER(C.c_int32))))
        h = C.c_void_p()
        _check(lib().sgpu_create(C.byref(model), C.byref(data), C.byref(h)))
        self._h = h
        self.fmt = fmt

    @classmethod
    def from_geno(cls, model: Model, geno: np.ndarray, nall: np.ndarray, fmt=None,
                  popflag: np.ndarray | None = None):
        f, words = pack(geno, fmt)
        return cls(model, f, words, nall, popflag)

    def get(self, which) -> np.ndarray:
        return self.get_chain(0, which)

    def set(self, which, value):
        self.set_chain(0, which, value)


class Batch(_Context):
    """Several chains advanced together (``sgpu_create_batch``): one launch per
    kernel and phase over all of them.

    ``datasets`` is a list of ``(fmt, words, nall)`` or ``(fmt, words, nall,
    popflag)``; chain c runs ``models[c]`` on ``datasets[data_of_chain[c]]``."""

    def __init__(self, models, datasets, data_of_chain=None, tuning=None):
        self.models = list(models)
        n = len(self.models)
        if data_of_chain is None:
            data_of_chain = [0] * n
        self._dsets = []
        darr = (Data * len(datasets))()
        for j, ds in enumerate(datasets):
            fmt, words, nall = ds[:3]
            pf = ds[3] if len(ds) > 3 else None
            w = np.ascontiguousarray(words, dtype=np.uint32)
            na = np.ascontiguousarray(nall, dtype=np.int32)
            pfa = None if pf is None else np.ascontiguousarray(pf, dtype=np.int32)
            self._dsets.append((w, na, pfa))
            darr[j] = Data(format=fmt, geno=w.ctypes.data_as(C.POINTER(C.c_uint32)),
                           geno_words=w.size, nall=na.ctypes.data_as(C.POINTER(C.c_int32)),
                           popflag=(None if pfa is None
                                    else pfa.ctypes.data_as(C.POINTER(C.c_int32))))
        marr = (Model * n)(*self.models)
        self._doc = np.ascontiguousarray(data_of_chain, dtype=np.int32)
        assert self._doc.size == n
        h = C.c_void_p()
        t = C.byref(tuning) if tuning is not None else None
        _check(lib().sgpu_create_batch(marr, n, darr, len(datasets), _vp(self._doc), t,
                                       C.byref(h)))
        self._h = h

    def __len__(self):
        return len(self.models)
