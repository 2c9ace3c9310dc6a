"""GPU 2D ADI (the reference's FluidSolver2D path) against the 2D oracle.

Bars: FP64 within 1e-12 relative L2 of the FP64 oracle on u, v, p and T, FP32
within 1e-5 of the FP32 and FP64 oracles, the same Picard iteration count per
step, and the divergence error to 1e-12 (FP64); the channel workload at full
size (256 x 128, 200 steps) and the long-line channel bit-identical in FP64
(the solve keeps the reference's operation order).
"""
import numpy as np
import pytest

import oracle_lib
from cmcx import domains
from cmcx.params import FluidParams
from cmcx.solver import AdiSolver2D
from gpu_util import rel_l2, say

pytestmark = pytest.mark.gpu


def _run(d, dtype, steps, G, L=1, params=None):
    p = params or FluidParams.normalized(d.re, 0.72, 1.4)
    s = AdiSolver2D(dtype).Init(d, p)
    o = oracle_lib.Oracle2D(d, dtype, params=p)
    its = []
    for i in range(steps):
        s.UpdateBoundaries()
        o.update_boundaries()
        e, it = s.TimeStep(d.dt, G, L)
        r, eo, ito = o.step(d.dt, G, L)
        assert r == 0
        its.append((it, ito, e, eo))
    return s, o, its, p


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("L", [1, 2])
def test_2d_matches_oracle(dtype, L):
    d = domains.channel2d(48, 24, radius=4.0, centre=(14, 12), dt=2e-3)
    say(f"[2d] {dtype} L={L}")
    s, o, its, p = _run(d, dtype, 6, 3, L)
    for it, ito, e, eo in its:
        assert it == ito
        if dtype == "f64":
            assert abs(e - eo) <= 1e-12 * max(abs(eo), 1e-30)
        else:
            assert abs(e - eo) <= 1e-4 * max(abs(eo), 1e-30)
    u, v, pp, T = s.fields()
    ou, ov, oT = o.layer(0)
    tol = 1e-12 if dtype == "f64" else 1e-5
    for nm, a, b in (("u", u, ou), ("v", v, ov), ("T", T, oT), ("p", pp, p.v_T * oT)):
        e = rel_l2(a, b)
        say(f"  {nm}: {e:.2e}")
        assert e < tol, nm
        if dtype == "f64" and nm != "p":
            assert np.array_equal(a, b), nm
    s.close()


def test_2d_channel_full_size():
    d = domains.channel2d()  # 256 x 128, r = 16 at (64, 64), Re 100, dt 1e-3, G 4
    steps = 200
    say(f"[2d channel] {d.nx}x{d.ny}, {steps} steps, FP64")
    s, o, its, p = _run(d, "f64", steps, d.num_global)
    say(f"  Picard iterations per step: min {min(i[0] for i in its)}, max {max(i[0] for i in its)}")
    assert all(i[0] == i[1] for i in its)
    u, v, pp, T = s.fields()
    ou, ov, oT = o.layer(0)
    for nm, a, b in (("u", u, ou), ("v", v, ov), ("T", T, oT)):
        e = rel_l2(a, b)
        say(f"  {nm}: rel L2 vs FP64 oracle {e:.2e}")
        assert np.array_equal(a, b), nm
    s.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("L", [1, 2])
def test_2d_long_lines(dtype, L):
    """X lines of 5000 rows: in FP64 their rows do not fit the LDS, so the X
    sweep runs the kernel with the global scratch, in FP32 they fit; the short
    Y lines take the LDS kernel in both.  L = 1 writes the direction's layers
    from the solve, L = 2 runs the copies and merges."""
    d = domains.channel2d(5000, 6, radius=1.5, centre=(2000, 3), dt=2e-3)
    say(f"[2d long lines] {d.nx}x{d.ny} {dtype} L={L}")
    s, o, its, p = _run(d, dtype, 3, 3, L)
    assert all(i[0] == i[1] for i in its)
    u, v, pp, T = s.fields()
    ou, ov, oT = o.layer(0)
    for nm, a, b in (("u", u, ou), ("v", v, ov), ("T", T, oT)):
        e = rel_l2(a, b)
        say(f"  {nm}: {e:.2e}")
        if dtype == "f64":
            assert np.array_equal(a, b), nm
        else:
            assert e < 1e-5, nm
    s.close()


@pytest.mark.parametrize("u_max, dt, at", [(1000.0, 2e-2, 1), (100.0, 2e-2, 2), (100.0, 2e-3, 7)])
def test_2d_picard_exit(u_max, dt, at):
    """A step whose divergence error passes 1.0 returns CMCX_ERR_PICARD at the
    oracle's iteration `at` with its error: in the first G = 3 iterations
    (queued back to back, their errors read at the end) or after them (one
    iteration at a time), and leaves the state as the oracle leaves it."""
    import ctypes
    from cmcx import _lib
    d = domains.channel2d(48, 24, radius=4.0, centre=(14, 12), dt=dt, u_max=u_max)
    p = FluidParams.normalized(d.re, 0.72, 1.4)
    say(f"[2d picard exit] u_max {u_max} dt {dt}")
    s = AdiSolver2D("f64").Init(d, p)
    o = oracle_lib.Oracle2D(d, "f64", params=p)
    s.UpdateBoundaries()
    o.update_boundaries()
    err = ctypes.c_double(0.0)
    it = ctypes.c_int(0)
    rc = _lib.lib().cmcx2d_step(s.h, ctypes.c_double(d.dt), 3, 1, ctypes.byref(err), ctypes.byref(it))
    ro, eo, ito = o.step(d.dt, 3, 1)
    say(f"  gpu rc {rc} it {it.value} err {err.value:.6e}; oracle rc {ro} it {ito} err {eo:.6e}")
    assert ro != 0 and ito == at
    assert rc == _lib.CMCX_ERR_PICARD
    assert it.value == ito
    assert abs(err.value - eo) <= 1e-12 * abs(eo)
    u, v, _, T = s.fields()
    ou, ov, oT = o.layer(0)
    for nm, a, b in (("u", u, ou), ("v", v, ov), ("T", T, oT)):
        assert np.array_equal(a, b), nm
    s.close()
