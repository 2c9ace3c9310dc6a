"""Pins of the 2D oracle (CPU, no GPU): the dense brute force written from the
model (tests/brute_adi2d.py) over several steps of random small domains,
known segment answers, and a closed form (viscous decay of a u mode)."""
import numpy as np
import pytest

import oracle_lib
from brute_adi2d import brute_step2d, segments2d
from cmcx import domains
from cmcx.params import FluidParams


def _random2d(nx, ny, seed, obstacle=True, free_outlet=True):
    rng = np.random.default_rng(seed)
    d = domains.channel2d(nx, ny, radius=min(nx, ny) / 5 if obstacle else 0.0,
                          centre=(nx / 3, ny / 2))
    t = d.type.reshape(nx, ny)
    if not free_outlet:
        d.bc.reshape(nx, ny)[-1, :] = domains.BC_NOSLIP
    # non-trivial inlet and wall data
    d.vy.reshape(nx, ny)[0, 1:-1] = 0.1 * rng.random(ny - 2)
    d.T.reshape(nx, ny)[0, :] = 1.0 + 0.1 * rng.random(ny)
    return d


@pytest.mark.parametrize("seed,G,L,obstacle,free", [(0, 1, 1, True, True), (1, 2, 1, True, False),
                                                    (2, 2, 2, False, True), (3, 3, 1, True, True)])
def test_step2d_vs_dense_brute_force(seed, G, L, obstacle, free):
    d = _random2d(11, 9, seed, obstacle, free)
    p = FluidParams(1.0, 0.05, 0.07, 0.2)
    dt = 2e-3
    o = oracle_lib.Oracle2D(d, "f64", params=p)
    state = [d.vx.copy(), d.vy.copy(), d.T.copy()]
    for step in range(3):
        o.update_boundaries()
        r, err, iters = o.step(dt, G, L)
        # the brute force starts from the boundary-filled state
        bv = (d.type == domains.NODE_BOUND) | (d.type == domains.NODE_VALVE)
        st = [np.where(bv, g, s) for g, s in zip((d.vx, d.vy, d.T), state)]
        rb, eb, ib, cur = brute_step2d(d, p, dt, G, L, st)
        assert r == rb == 0 and iters == ib
        assert abs(err - eb) <= 1e-12 * max(abs(eb), 1e-30)
        got = o.layer(0)
        for k in range(3):
            assert np.allclose(got[k], cur[k].ravel(), rtol=1e-11, atol=1e-12), (step, k)
        state = [c.ravel() for c in cur]


def test_segments2d_known_answers():
    d = domains.channel2d(12, 8, radius=2.0, centre=(5, 4))
    o = oracle_lib.Oracle2D(d, "f64")
    rows = o.segments(0)   # X-direction systems, one per fluid row
    cols = o.segments(1)   # Y-direction systems, one per column with fluid
    br, bc = segments2d(d.type.reshape(12, 8))
    assert [tuple(r[[0, 1, 2]]) for r in rows] == [(a, j, b) for (a, j, b) in br]
    assert [(c[0], c[1], c[3]) for c in cols] == [(i, a, b) for (i, a, b) in bc]
    # channel: every interior row spans inlet to outlet, every interior column wall to wall
    assert len(rows) == 6 and all(r[0] == 0 and r[2] == 11 and r[4] == 12 for r in rows)
    assert len(cols) == 10 and all(c[1] == 0 and c[3] == 7 and c[4] == 8 for c in cols)


def test_viscous_decay_of_u_mode():
    """u = A sin(pi y) uniform in x, v = 0, T = 1, v_T = t_phi = 0, walls at
    y = 0 and y = 1 (no-slip), BC_FREE ends in x.  The X sweep leaves u exactly
    unchanged (a uniform profile solves its rows for any advecting u and
    satisfies x0 = x1); the Y sweep is a pure diffusion solve (v = 0), so one
    Picard iteration multiplies the mode by 1 / (1 + dt nu lambda) with
    lambda = (4/h^2) sin^2(pi h / 2), the eigenvalue of the 3-point Laplacian
    with Dirichlet ends."""
    nx, ny = 7, 17
    d = domains.channel2d(nx, ny, radius=0.0)
    t = d.type.reshape(nx, ny)
    t[:] = domains.NODE_IN
    t[0, :] = t[-1, :] = domains.NODE_VALVE
    t[:, 0] = t[:, -1] = domains.NODE_BOUND
    bc = d.bc.reshape(nx, ny)
    bc[:] = domains.BC_NOSLIP
    bc[0, 1:-1] = bc[-1, 1:-1] = domains.BC_FREE
    h = 1.0 / (ny - 1)
    d.dx = d.dy = h
    A = 0.3
    u0 = np.broadcast_to(A * np.sin(np.pi * np.arange(ny) * h)[None, :], (nx, ny)).copy()
    d.vx = u0.ravel().copy()
    d.vy = np.zeros(nx * ny)
    d.T = np.ones(nx * ny)
    nu, dt = 0.1, 1e-3
    p = FluidParams(0.0, nu, nu, 0.0)
    o = oracle_lib.Oracle2D(d, "f64", params=p)
    o.update_boundaries()
    r, err, iters = o.step(dt, 1, 1)
    assert r == 0 and iters == 1 and err < 1e-12
    u = o.layer(0)[0].reshape(nx, ny)
    lam = 4.0 / h**2 * np.sin(np.pi * h / 2) ** 2
    want = u0 / (1.0 + dt * nu * lam)
    assert np.allclose(u[1:-1, 1:-1], want[1:-1, 1:-1], rtol=1e-12, atol=0)
