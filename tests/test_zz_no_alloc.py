"""GPU test of the allocation-free entry point SimInf_hip_abc_weights_noalloc:
through the C ABI, its weights equal SimInf_hip_abc_weights' bit for bit
(and its error codes), on consecutive calls of which the second is
smaller (the fixed device array keeps the first call's contents), and
it refuses sizes beyond its array."""
import numpy as np
import pytest

from siminf_amd import _native
from siminf_amd import abc_smc as A

pytestmark = pytest.mark.gpu


def _case(n, d, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.2, 0.8, size=(n, d))
    xx = np.clip(x + rng.normal(0, 0.05, size=(n, d)), 0.01, 0.99)
    w = rng.uniform(0.5, 1.5, size=n)
    w /= w.sum()
    sigma = np.diag(np.full(d, 0.01)) + 0.001
    dist = "ugn"[:d] if d <= 3 else "u" * d
    p1 = np.array([0.0, 2.0, 0.5][:d] + [0.0] * max(0, d - 3))
    p2 = np.array([1.0, 5.0, 0.2][:d] + [1.0] * max(0, d - 3))
    return dist, p1, p2, x, xx, w, sigma


def _both(lib, case):
    dist, p1, p2, x, xx, w, sigma = case
    a = A.call_weights(lib.SimInf_hip_abc_weights, dist, p1, p2, x, xx, w, sigma)
    b = A.call_weights(lib.SimInf_hip_abc_weights_noalloc, dist, p1, p2, x, xx, w, sigma)
    return a, b


def test_noalloc_equals_allocating_entry_point():
    lib = _native.load()
    for n, d, seed in ((3000, 3, 1), (700, 2, 2), (5, 1, 3)):
        print("case n=%d d=%d seed=%d" % (n, d, seed), flush=True)
        (ea, wa), (eb, wb) = _both(lib, _case(n, d, seed))
        assert ea == eb == 0
        np.testing.assert_array_equal(wa, wb)
        assert abs(wb.sum() - 1.0) < 1e-12


def test_noalloc_errors_and_limits():
    lib = _native.load()
    # A prior density of 0 (a proposal outside a uniform prior's bounds).
    dist, p1, p2, x, xx, w, sigma = _case(200, 1, 4)
    xx[17, 0] = 2.0
    (ea, _), (eb, _) = _both(lib, (dist, p1, p2, x, xx, w, sigma))
    assert ea == eb == -102
    # A following smaller call starts from a clean error word.
    (ea, wa), (eb, wb) = _both(lib, _case(50, 1, 5))
    assert ea == eb == 0
    np.testing.assert_array_equal(wa, wb)
    # Without previous weights every weight is 1/n.
    e, ww = A.call_weights(lib.SimInf_hip_abc_weights_noalloc, "u", None, None, None,
                           np.zeros((8, 1)), None, None)
    assert e == 0 and (ww == 1 / 8).all()
    # Beyond the fixed array: about 1.1e6 particles of 3 parameters.
    dist, p1, p2, x, xx, w, sigma = _case(1_100_000, 3, 6)
    e, _ = A.call_weights(lib.SimInf_hip_abc_weights_noalloc, dist, p1, p2, x, xx, w, sigma)
    assert e == -2
