"""cmcx_div_error_fields, the entry point that allocates nothing (its partial
sums live in a fixed device array of the library), against cmcx_div_error,
the entry point it is a variant of, through the C ABI: bit-identical errors on
two calls in a row, the second on a smaller grid, in FP32 and FP64."""
import ctypes

import numpy as np
import pytest

from cmcx import _lib, domains
from cmcx.solver import make_solver
from gpu_util import say

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        _lib.lib()
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipFree.argtypes = [ctypes.c_void_p]
        self.h = h
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.bufs.append(p.value)
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _case(d, dtype, steps, hip):
    s = make_solver(d, dtype, path=1)
    for _ in range(steps):
        s.UpdateBoundaries()
        s.TimeStep(d.dt, 2, 1, False)
    want = s.div_error()  # cmcx_div_error on the current state
    nb = d.n * np.dtype(s.np_dtype).itemsize
    u, v, w = (hip.alloc(nb) for _ in range(3))
    s.fields_device(u, v, w, None, None)
    s.synchronize()
    t = hip.upload(np.ascontiguousarray(d.type, dtype=np.uint8))
    dims = (ctypes.c_int32 * 3)(d.nx, d.ny, d.nz)
    h = (ctypes.c_double * 3)(d.dx, d.dy, d.dz)
    got = ctypes.c_double(0.0)
    rc = _lib.lib().cmcx_div_error_fields(u, v, w, t, dims, h, s.dtype, None, ctypes.byref(got))
    assert rc == 0, _lib.lib().cmcx_last_error()
    s.close()
    return got.value, want


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_div_error_fields_matches_solver(dtype):
    hip = _Hip()
    try:
        big = domains.ellipsoids(48, count=6, axis_lo=3, axis_hi=9, seed=2, dt=5e-4)
        small = domains.lid_sphere(20, radius=4.0, dt=1e-3)
        for d, steps in ((big, 3), (small, 2)):  # the second call on a smaller grid
            got, want = _case(d, dtype, steps, hip)
            say(f"[no-alloc div error] {dtype} {d.shape}: {got!r} vs {want!r}")
            assert np.isfinite(want) and want > 0
            assert got == want
    finally:
        hip.free()


def test_div_error_fields_rejects_bad_arguments():
    L = _lib.lib()
    dims = (ctypes.c_int32 * 3)(2, 8, 8)
    h = (ctypes.c_double * 3)(1.0, 1.0, 1.0)
    e = ctypes.c_double(0.0)
    # null pointers and too-small grids are refused before any device work
    assert L.cmcx_div_error_fields(None, None, None, None, dims, h, 0, None, ctypes.byref(e)) == _lib.CMCX_ERR_ARG
    one = ctypes.c_void_p(1)
    assert L.cmcx_div_error_fields(one, one, one, one, dims, h, 0, None, ctypes.byref(e)) == _lib.CMCX_ERR_ARG
    big = (ctypes.c_int32 * 3)(2048, 2048, 1024)
    assert L.cmcx_div_error_fields(one, one, one, one, big, h, 0, None, ctypes.byref(e)) == _lib.CMCX_ERR_TOO_LARGE
