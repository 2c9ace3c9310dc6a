"""GPU 2D ADI (the reference's FluidSolver2D path) against the 2D oracle.

Bars: FP64 within 1e-12 relative L2 of the FP64 oracle on u, v, p and T, FP32
within 1e-5 of the FP32 and FP64 oracles, the same Picard iteration count per
step, and the divergence error to 1e-12 (FP64); the channel workload at full
size (256 x 128, 200 steps) within 1e-11.
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
        assert e < 1e-11, nm
    s.close()
