"""The allocation-free continuity residual (pcp_continuity_residual_noalloc)
against pcp_continuity_residual, through the C ABI, on two calls in a row:
first a 32 x 64 grid over 2 x 2 domains, then a smaller 16 x 32 grid on one
domain, so the second call runs over what the first left in the fixed
array.  Villasenor-Buneman deposition in FP64 conserves charge, so both
residuals stay within 1e-12 of the largest macro-particle charge; the
residuals agree to that level (the charge sums are atomics), the charge
change to 1e-12 of itself, the largest |q w| exactly.  A grid beyond the
fixed array is refused."""
import ctypes as C

import numpy as np

import pytest

from picopic_amd import capi
from tests.test_gpu_continuity import make

pytestmark = pytest.mark.gpu


def residuals(ctx):
    L = capi.load()
    L.pcp_continuity_residual_noalloc.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    a, b = np.zeros(3), np.zeros(3)
    capi._check(L.pcp_continuity_residual(ctx.h, a.ctypes.data_as(C.POINTER(C.c_double))))
    capi._check(L.pcp_continuity_residual_noalloc(ctx.h, b.ctypes.data_as(C.POINTER(C.c_double))))
    return a, b


def test_noalloc_residual_matches_on_two_calls():
    for domains, n, cells in (((2, 2), 6000, (32, 64)), ((1, 1), 2000, (16, 32))):
        ctx = make("vb", "fp64", domains, n=n, seed=n, cells=cells)
        ctx.step(2)
        a, b = residuals(ctx)
        print("grid", cells, "domains", domains, "residual", a, b, flush=True)
        assert a[1] > 0 and abs(a[1] - b[1]) <= 1e-12 * a[1]
        assert a[2] == b[2]
        assert a[0] <= 1e-12 * a[2] and b[0] <= 1e-12 * b[2]
        assert abs(a[0] - b[0]) <= 1e-12 * a[2]
        ctx.close()


def test_noalloc_refuses_a_grid_beyond_the_array():
    ctx = capi.Context((4.0, 4.0), (2048, 8192), (1, 1), 1e-12, [(-1, 1, False, 16)])
    L = capi.load()
    L.pcp_continuity_residual_noalloc.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    out = np.zeros(3)
    rc = L.pcp_continuity_residual_noalloc(ctx.h, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc != 0 and b"exceeds the fixed scratch array" in L.pcp_last_error()
    ctx.close()
