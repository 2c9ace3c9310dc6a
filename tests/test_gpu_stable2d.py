"""The Stable 2D solver on the GPU (csrc/solver2d.hip, CMCX2D_STABLE) against
the oracle's restatement of StableSolver2D (oracle/oracle2d.cpp), same
precision, bit for bit: states after whole steps and the relaxation's sweep
count of every project.  Random walled grids (obstacles, outside regions,
boundary cells on the grid's edge), the reference's four 2D example inputs
end to end (grid built on the GPU, Grid2D::Prepare at each step's time, the
mask changing between frames), a grid too large for LDS (the relaxation's
global-memory variant), and the error paths."""
import os

import numpy as np
import pytest

import oracle_lib
from cmcx._lib import CmcxError
from cmcx.domains import Domain2D
from cmcx.geometry import Grid2D
from cmcx.params import FluidParams
from cmcx.solver import StableSolver2D
from gpu_util import say
from test_stable2d_pins import domain, walled

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD2 = os.path.join(HERE, "golden", "geometry2d")

IN, OUT, BOUND, VALVE = 0, 1, 2, 3


def _pair(d, dtype, p):
    o = oracle_lib.Oracle2D(d, dtype, p, method="stable")
    s = StableSolver2D(dtype).Init(d, p)
    return o, s


def _same_state(s, o, what):
    u, v, _, T = s.fields()
    ru, rv, rT = o.layer(0)
    for nm, a, b in (("u", u, ru), ("v", v, rv), ("T", T, rT)):
        nd = int(np.count_nonzero(a.astype(np.float64) != b))
        assert nd == 0, f"{what}: {nm} differs at {nd} nodes (max {np.abs(a - b).max():.3e})"


def _steps(s, o, dt, G, n, what, grids=None):
    for k in range(n):
        if grids is not None:
            arrs = grids(k)
            s.SetGrid(*arrs)
            o.set_grid(*arrs)
        s.UpdateBoundaries()
        o.update_boundaries()
        err, it = s.TimeStep(dt, G, 1)
        r, rerr, rit = o.step(dt, G, 1)
        assert r == 0
        sw, rsw = list(s.poisson_sweeps()), list(o.poisson_sweeps())
        say(f"  {what} step {k}: iters {it}, sweeps {sw}, err {err:.3e}")
        assert it == rit
        assert sw == rsw
        np.testing.assert_allclose(err, rerr, rtol=1e-5)
        _same_state(s, o, f"{what} step {k}")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_walled_grids_bit_exact(dtype, seed):
    nx, ny = [(40, 33), (64, 70), (97, 50)][seed]
    t = walled(nx, ny, seed, obstacles=12)
    if seed == 1:
        t[0, 5] = BOUND
        t[nx - 1, 7] = VALVE
    d = domain(t, 0.01, 0.012, seed=seed)
    p = FluidParams(1.0, 0.02, 0.0, 0.0)
    o, s = _pair(d, dtype, p)
    say(f"[stable2d] walled {nx}x{ny} {dtype} seed {seed}")
    _steps(s, o, 2e-4, 2, 3, "walled")
    s.close()


EXAMPLES = {
    # file, dx, viscosity, steps per frame (the reference's configs)
    "box_pipe": ("box_pipe_data.txt", 0.0007, 0.05, 30),
    "heart_MR": ("heart_MR_data.txt", 0.0007, 0.05, 30),
    "heart_US": ("heart_US_data.txt", 0.0007, 0.05, 50),
    "white_sea": ("white_sea_data.txt", 0.002, 0.01, 5000),
}


@pytest.mark.parametrize("name", list(EXAMPLES))
def test_example_grids_bit_exact(name):
    fn, dx, _, _ = EXAMPLES[name]
    path = os.path.join(GOLD2, fn)
    g = Grid2D(path, dx, dx, startT=1.0)
    og = oracle_lib.OracleGeometry("shape2d", path, dx, dx, dx, depth=0.0)
    times = [0.0, 0.3 * g.cycle_length / g.num_frames, 0.77 * g.cycle_length]
    for time in times:
        got = g.arrays(time)
        want = og.build2d(time)
        assert g.shape == og.dims2d
        for nm, a, b in zip(("type", "bc", "vx", "vy", "T"), got, want):
            nd = int(np.count_nonzero(a != b))
            assert nd == 0, f"{name} t={time}: {nm} differs at {nd} cells"
    say(f"[stable2d] {name} grid {g.shape}, {g.num_frames} frames")
    g.close()


@pytest.mark.parametrize("name,dtype", [("box_pipe", "f64"), ("heart_MR", "f64"), ("heart_US", "f32"),
                                        ("white_sea", "f32"), ("heart_MR", "f32")])
def test_example_runs_end_to_end(name, dtype):
    """The reference's 2D driver loop (FluidSolver2D.cpp:110-135): Prepare(t),
    UpdateBoundaries, TimeStep(dt, num_global 2, num_local 1), at times that
    cross frames so the walls (and the mask) move between steps."""
    fn, dx, vis, spf = EXAMPLES[name]
    path = os.path.join(GOLD2, fn)
    g = Grid2D(path, dx, dx, startT=1.0)
    dt = g.cycle_length / (g.num_frames * spf)
    frame = g.cycle_length / g.num_frames
    times = [dt, 0.5 * frame + dt, 1.5 * frame + dt]
    d = g.Prepare(0.0)
    p = FluidParams.dimensional(vis, 1000.0, 287.0, 0.0257, 1000.0)
    o, s = _pair(d, dtype, p)
    say(f"[stable2d] {name} {dtype} {g.shape}, dt {dt:.3e}")
    _steps(s, o, dt, 2, len(times), name, grids=lambda k: g.arrays(times[k]))
    s.close()
    g.close()


def test_global_memory_relaxation_bit_exact():
    """A grid whose q and flags exceed LDS: the relaxation keeps them in
    global memory (the kernel's second variant)."""
    nx, ny = 300, 200
    t = walled(nx, ny, 9, obstacles=40)
    d = domain(t, 0.004, 0.005, seed=9)
    p = FluidParams(1.0, 0.01, 0.0, 0.0)
    o, s = _pair(d, "f64", p)
    o.set_poisson(0.05, 1 << 30)
    s.set_poisson(0.05)
    say(f"[stable2d] global-memory variant {nx}x{ny}")
    _steps(s, o, 1e-4, 1, 2, "global")
    s.close()


def test_errors():
    p = FluidParams(1.0, 0.02, 0.0, 0.0)
    t = walled(20, 16, 0)
    t[0, 5] = IN  # fluid on the grid's edge
    with pytest.raises(CmcxError):
        StableSolver2D("f64").Init(domain(t, 0.01, 0.01), p)
    big = np.full((1100, 8), BOUND, np.uint8)
    big[1:-1, 1:-1] = IN
    with pytest.raises(CmcxError):
        StableSolver2D("f64").Init(domain(big, 0.01, 0.01), p)
    # a sweep cap that the relaxation reaches: the step fails as the reference's exits do
    t = walled(30, 24, 1)
    s = StableSolver2D("f64").Init(domain(t, 0.01, 0.01, seed=1), p)
    s.set_poisson(1e-12, 3)
    s.UpdateBoundaries()
    with pytest.raises(CmcxError):
        s.TimeStep(1e-3, 2, 1)
    with pytest.raises(CmcxError):
        s.set_poisson(0.0, 10)
    s.close()


def test_still_fluid_stays_still():
    """No motion anywhere: every project relaxes a zero divergence, each sweep
    ends on a 0/0 change, one sweep per project; the state stays zero."""
    t = walled(24, 20, 2)
    d = domain(t, 0.01, 0.01)
    d.vx[:] = 0
    d.vy[:] = 0
    p = FluidParams(1.0, 0.02, 0.0, 0.0)
    s = StableSolver2D("f64").Init(d, p)
    for _ in range(2):
        s.UpdateBoundaries()
        s.TimeStep(1e-3, 2, 1)
        assert list(s.poisson_sweeps()) == [1, 1]
    u, v, _, T = s.fields()
    assert not u.any() and not v.any()
    s.close()
