"""sgpu_eval_math_noalloc, the allocation-free variant of sgpu_eval_math: the
same device functions on the same inputs give the same values bit for bit,
over two calls in a row on the shared static staging (the second smaller, so
stale values of the first would show), and sizes beyond the staging are
refused without touching the device."""
import numpy as np
import pytest

import structure_amd as sa

pytestmark = pytest.mark.gpu


def _inputs(fn, n, seed):
    rng = np.random.default_rng(seed)
    if fn == sa.FN_LOG:
        return np.exp(rng.uniform(-700, 700, n))
    if fn == sa.FN_LGAMMA:
        return np.exp(rng.uniform(-5, 12, n))
    return rng.uniform(0, 8, n)


@pytest.mark.parametrize("fn", [sa.FN_LOG, sa.FN_COS, sa.FN_SIN, sa.FN_LGAMMA])
def test_noalloc_equals_allocating_entry_point(fn):
    for n, seed in ((sa.EVAL_MATH_NOALLOC_MAX, 1), (1537, 2)):
        x = _inputs(fn, n, seed + 10 * fn)
        want = sa.eval_math(fn, x)
        got = sa.eval_math(fn, x, noalloc=True)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (fn, n)


def test_noalloc_refuses_sizes_beyond_its_staging():
    x = np.ones(sa.EVAL_MATH_NOALLOC_MAX + 1)
    with pytest.raises(sa.SgpuError) as e:
        sa.eval_math(sa.FN_LOG, x, noalloc=True)
    assert e.value.code == sa.ERR_ARG
    assert sa.eval_math(sa.FN_LOG, np.ones(0), noalloc=True).size == 0
    y = sa.eval_math(sa.FN_LOG, np.array([1.0, 2.0]), noalloc=True)   # still usable after
    assert y[0] == 0.0 and abs(y[1] - np.log(2.0)) < 1e-15
