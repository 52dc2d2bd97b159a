"""qlx_raytrace_noalloc (device staging in a file-scope array, no allocation) against qlx_raytrace,
through the C ABI: bit-identical results on two calls in a row, the second with a smaller input,
and a clear error beyond the array's capacity."""
import ctypes as C

import numpy as np
import pytest

from qlens_amd import qlx
from qlens_amd import workloads as W

pytestmark = pytest.mark.gpu


def _call(inv, name, lenses, xy):
    xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
    out = np.zeros_like(xy)
    m = qlx.make_model(lenses)
    fn = getattr(inv.lib, name)
    fn.restype = C.c_int
    rc = fn(inv.ctx, C.byref(m), len(xy), xy.ctypes.data_as(C.POINTER(C.c_double)), out.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, out


def test_noalloc_raytrace_equals_raytrace():
    inv = qlx.PixelInversion(0)
    try:
        case = W.c2()
        inv.set_settings(2, (-1, 1, -1, 1), (10, 10), 1e3, 1e-6)
        rng = np.random.default_rng(17)
        for n in (200000, 3001):  # the second call smaller than the first
            xy = rng.uniform(-4, 4, size=(n, 2))
            rc0, ref = _call(inv, "qlx_raytrace", case["lenses"], xy)
            rc1, got = _call(inv, "qlx_raytrace_noalloc", case["lenses"], xy)
            print("n %d: max |d beta| %.3g" % (n, np.abs(got - ref).max()), flush=True)
            assert rc0 == 0 and rc1 == 0
            assert np.array_equal(got, ref)
        rc, _ = _call(inv, "qlx_raytrace_noalloc", case["lenses"], np.zeros((262145, 2)))
        assert rc == -7  # QLX_ERR_RANGE, nothing launched
    finally:
        inv.close()
