"""The 128^3 lid-driven box with a sphere at full size (100 steps, 4 Picard
iterations, Re 400, dt 5e-4) on the GPU, FP32 and FP64, fused and plain,
against one FP64 oracle run of the same 100 steps.

Bars: FP64 within 1e-11 and FP32 within 1e-5 relative L2 of the FP64 oracle
on u, v, w (velocity-norm scaled) and p = v_T T.  The errors are printed.
"""
import numpy as np
import pytest

import oracle_lib
from cmcx import domains
from cmcx.solver import make_solver
from gpu_util import rel_l2, rel_l2_scaled, say, vel_norm

pytestmark = pytest.mark.gpu

STEPS = 100


@pytest.fixture(scope="module")
def case():
    d = domains.lid_sphere(128)
    say(f"[C2] oracle FP64, {STEPS} steps of {d.shape} (CPU)")
    o = oracle_lib.Oracle3D(d, "f64")
    o.create_segments()
    r, _ = o.run(STEPS, d.dt, 4, 1)
    assert r == 0
    f = o.fields()
    o.close()
    return d, f


@pytest.mark.parametrize("dtype,path", [("f64", 1), ("f32", 1), ("f32", 0), ("f64", 0)])
def test_c2_full_size(case, dtype, path):
    d, want = case
    s = make_solver(d, dtype, path=path)
    s.run(STEPS, d.dt, 4, 1)
    u, v, w, p, T = s.fields()
    vs = vel_norm(*want[:3])
    tol = 1e-11 if dtype == "f64" else 1e-5
    for nm, a, b, sc in (("u", u, want[0], vs), ("v", v, want[1], vs), ("w", w, want[2], vs),
                         ("p", p, want[3] * d.params().v_T, 0.0)):
        e = rel_l2_scaled(a, b, sc)
        say(f"  C2 {dtype} path={path} {nm}: rel L2 vs FP64 oracle {e:.3e}")
        assert e < tol, nm
    s.close()
