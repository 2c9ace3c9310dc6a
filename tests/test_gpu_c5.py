"""The gradient workload at its configured size: a 96^3 channel with a seeded
z-axis cylinder (radius 12), inlet driven by a 32x32 control (each value fills
a 3x3 block of inlet nodes), nu = 1e-3, 200 steps of 4 Picard iterations,
checkpoint every 20 steps; loss = mean kinetic energy over a 16^3 probe
behind the cylinder; gradient with respect to the control and nu.

Bars: the FP64 and the FP32 adjoints' directional derivatives within 1e-4 of
central differences (eps 1e-3) of the FP64 forward along 4 seeded random
directions in (control, nu); the FP32 adjoint's gradient within 1e-3 relative
L2 of the FP64 adjoint's (control) and 1e-3 relative (nu).
"""
import numpy as np
import pytest

import adjoint_cases as ac
from cmcx.solver import make_solver
from gpu_util import rel_l2, say

pytestmark = pytest.mark.gpu

EVERY = 20
NDIR = 4
EPS = 1e-3


@pytest.fixture(scope="module")
def problem():
    return ac.ChannelProblem(n=96, steps=200)


def _gradient(pb_, dtype):
    d = pb_.d
    s = make_solver(d, dtype, path=1)
    init = s.fields()
    s.adjoint_forward(pb_.steps, d.dt, d.num_global, EVERY)
    loss, seed = pb_.loss(s.fields())
    gb, _, pp = s.adjoint_backward(seed)
    s.adjoint_release()
    return s, loss, pb_.block_sum(gb[0]), float(pp[1]), [init[0], init[1], init[2], init[4]]


def test_c5_gradient(problem):
    pb_ = problem
    d = pb_.d
    say(f"[C5] {d.shape} channel, {pb_.steps} steps, checkpoint every {EVERY}")
    s64, L64, gc64, gnu64, st0 = _gradient(pb_, "f64")
    s32, L32, gc32, gnu32, _ = _gradient(pb_, "f32")
    s32.close()
    say(f"  loss FP64 {L64:.10e}  FP32 {L32:.10e}")
    e_ctl = rel_l2(gc32, gc64)
    e_nu = abs(gnu32 - gnu64) / abs(gnu64)
    say(f"  FP32 vs FP64 adjoint: control gradient {e_ctl:.2e}, nu {e_nu:.2e}")
    assert e_ctl < 1e-3 and e_nu < 1e-3
    nu0 = d.params().v_vis

    def L(c, nu):
        p = d.params()
        p.v_vis = nu
        s64.set_params(p)
        s64.set_boundary(d.bc_vel, d.bc_temp, pb_.inlet(c), d.vy, d.vz, d.T)
        s64.set_fields(*st0)
        for _ in range(pb_.steps):
            s64.UpdateBoundaries()
            s64.TimeStep(d.dt, d.num_global, 1, False)
        return pb_.loss(s64.fields())[0]

    rng = np.random.default_rng(0)
    for k in range(NDIR):
        dc = rng.normal(size=pb_.control.shape)
        dn = rng.normal() * nu0
        fd = (L(pb_.control + EPS * dc, nu0 + EPS * dn)
              - L(pb_.control - EPS * dc, nu0 - EPS * dn)) / (2 * EPS)
        a64 = float(np.sum(gc64 * dc) + gnu64 * dn)
        a32 = float(np.sum(gc32 * dc) + gnu32 * dn)
        r64, r32 = abs(a64 - fd) / abs(fd), abs(a32 - fd) / abs(fd)
        say(f"  direction {k}: FD {fd:.10e}  FP64 adjoint {a64:.10e} ({r64:.2e})  "
            f"FP32 adjoint {a32:.10e} ({r32:.2e})")
        assert r64 < 1e-4 and r32 < 1e-4
    s64.close()
