"""adg_sweep_noalloc (the allocation-free variant of adg_sweep, staging in the
library's per-GPU scratch array) against adg_sweep, through the C-ABI: two
calls in a row, the second on a smaller input, bit-identical outputs; sizes
beyond the scratch capacity are refused before any GPU work."""
import ctypes

import numpy as np
import pytest

import pyadg

pytestmark = pytest.mark.gpu

I32P = ctypes.POINTER(ctypes.c_int32)
DP = ctypes.POINTER(ctypes.c_double)


def _noalloc():
    f = pyadg.lib().adg_sweep_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, I32P, I32P,
                  ctypes.c_double, DP, DP]
    return f


def _call(fn, s, table, flx, lu, d, si, so, coef, x):
    out = np.zeros(s.num_elem * int(np.prod(so)))
    a = (ctypes.c_int32 * len(si))(*si)
    b = (ctypes.c_int32 * len(so))(*so)
    rc = fn(s._h, table, flx, lu, d, a, b, coef, x.ctypes.data_as(DP), out.ctypes.data_as(DP))
    return rc, out


def test_sweep_noalloc_matches_sweep_twice():
    print("[case] adg_sweep_noalloc vs adg_sweep, two calls", flush=True)
    fn = _noalloc()
    s = pyadg.Solution(2, 2, 3, 7, True, 7)
    rng = np.random.default_rng(3)
    cases = [(3, 1, 0, 0, [4, 4], [3, 4], 0.5), (0, 0, 2, 1, [3, 3], [3, 4], -1.25)]
    for n, (table, flx, lu, d, si, so, coef) in enumerate(cases):
        if n == 1:
            s = pyadg.Solution(2, 2, 3, 7, True, 5)  # the second call: a smaller grid
        x = rng.standard_normal(s.num_elem * int(np.prod(si)))
        rc, got = _call(fn, s, table, flx, lu, d, si, so, coef, x)
        assert rc == 0, pyadg.lib().adg_last_error(s._h)
        want = s.sweep(table, flx, lu, d, si, so, coef, x)
        assert np.array_equal(got, want), f"call {n}: max abs diff {np.max(np.abs(got - want))}"


def test_sweep_noalloc_refuses_oversize():
    fn = _noalloc()
    s = pyadg.Solution(2, 3, 4, 9, False, 9)  # full grid: 2^18 elements x (16 + 20) values > 2^22
    si, so = [5, 5], [4, 5]
    x = np.zeros(s.num_elem * 25)
    rc, _ = _call(fn, s, 3, 1, 0, 0, si, so, 1.0, x)
    assert rc == -1 and b"scratch capacity" in pyadg.lib().adg_last_error(s._h)
