"""The device special functions of the sampler against the host's libm.

log_d and sincos_d are the sampler's own (atanh series / Taylor series with
Cody-Waite reduction); they must stay within 2 ulp of glibc, whose log, cos
and sin are within an ulp of the true values.  lgamma_lanczos is the reference's
Lanczos sum (mymath.c:382-429): against the oracle's term-by-term restatement
within a few ulps, and against math.lgamma within the approximation's own
error
(about 6e-12 relative for z below 1e-4).
"""
import math

import numpy as np
import pytest

import oracle
import structure_amd as sa

pytestmark = pytest.mark.gpu


def _ulps(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    u = np.spacing(np.abs(b))
    return np.where(a == b, 0.0, np.abs(a - b) / u)


def test_log_within_two_ulps():
    rng = np.random.default_rng(1)
    x = np.concatenate([
        np.exp(rng.uniform(-744, 709, 400_000)),          # whole range incl. subnormals
        1.0 + rng.uniform(-0.4, 0.4, 200_000),           # around 1, where ln x is small
        (rng.integers(0, 2**32, 200_000, dtype=np.uint64) + 0.5) * 2.0**-32,   # uniforms
        [5e-324, 2.2250738585072014e-308, 0.7071067811865475, 0.7071067811865476, 1.0,
         2.0, 1.7976931348623157e308],
    ])
    y = sa.eval_math(sa.FN_LOG, x)
    ref = np.log(x)
    assert _ulps(y, ref).max() <= 2.0
    assert y[np.flatnonzero(x == 1.0)[0]] == 0.0
    sp = sa.eval_math(sa.FN_LOG, np.array([0.0, np.inf, -1.0, np.nan, -0.0]))
    assert sp[0] == -np.inf and sp[1] == np.inf and sp[4] == -np.inf
    assert np.isnan(sp[2]) and np.isnan(sp[3])


def test_cos_and_sin_of_box_muller_angles():
    rng = np.random.default_rng(2)
    w = np.concatenate([
        rng.integers(0, 2**32, 300_000, dtype=np.uint64),
        # the quarter turns, where cos crosses zero or peaks
        *[np.arange(c - 2000, c + 2000, dtype=np.uint64) for c in (2**30, 2**31, 3 * 2**30)],
        np.array([0, 2**32 - 1], dtype=np.uint64),
    ])
    a = 6.283185307179586 * ((w.astype(np.float64) + 0.5) * 2.0**-32)
    y = sa.eval_math(sa.FN_COS, a)
    ref = np.array([math.cos(v) for v in a])
    assert _ulps(y, ref).max() <= 2.0
    y = sa.eval_math(sa.FN_SIN, a)
    ref = np.array([math.sin(v) for v in a])
    assert _ulps(y, ref).max() <= 2.0


def test_lgamma_matches_lanczos_and_libm():
    rng = np.random.default_rng(3)
    z = np.concatenate([np.exp(rng.uniform(-12, 14, 20_000)), [0.5, 1.0, 2.0, 3.5, 1e6, 1e35]])
    y = sa.eval_math(sa.FN_LGAMMA, z)
    ora = np.array([oracle.lgamma(v) for v in z])
    scale = np.maximum(np.abs(ora), 1.0)
    assert (np.abs(y - ora) / scale).max() < 1e-14
    # the Lanczos sum itself departs from ln Gamma by up to about 6e-12 for z
    # below 1e-4 (same in the oracle), and by about 1e-15 elsewhere
    lib = np.array([math.lgamma(v) for v in z])
    err = np.abs(y - lib) / np.maximum(np.abs(lib), 1.0)
    assert err.max() < 2e-11
    assert err[z > 1e-2].max() < 1e-13


def test_eval_math_rejects_bad_arguments():
    lib = sa.lib()
    assert lib.sgpu_eval_math(7, None, None, 1) == sa.ERR_ARG
    assert lib.sgpu_eval_math(sa.FN_LOG, None, None, -1) == sa.ERR_ARG
