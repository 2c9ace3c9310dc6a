"""Pins of the oracle's Stable 2D solver (oracle/oracle2d.cpp, restating
src/FluidSolver2D/StableSolver2D.cpp:21-234) on the CPU:

  * SolveU / SolveV against the closed form for quadratic fields (central
    differences are exact there), with the state and the Picard layer
    different and dx != dy, so a swapped layer, axis, sign or dropped term
    fails;
  * the relaxation against a matrix-splitting model (tests/brute_stable2d.py:
    each sweep a triangular solve in the sweep's cell order), after 1, 2 and 7
    sweeps, on masks with obstacles, outside regions and boundary cells on the
    grid's edge;
  * the stopping test (the std::max fold in which a 0/0 change resets the
    running value) against the model's vectorised fold: sweep counts equal,
    and a zero field stops after one sweep;
  * whole steps (copies, merges, Picard loop, ClearOutterCells) against the
    model's masked-array restatement.
"""
import numpy as np
import pytest

import brute_stable2d as B
import oracle_lib
from cmcx.domains import Domain2D
from cmcx.params import FluidParams

IN, OUT, BOUND, VALVE = 0, 1, 2, 3


def walled(nx, ny, seed, out_corner=True, obstacles=4):
    """Walls on the border (a valve column on the left), an outside corner
    region behind a wall, random obstacle cells."""
    rng = np.random.default_rng(seed)
    t = np.full((nx, ny), IN, np.uint8)
    t[0, :] = VALVE
    t[-1, :] = BOUND
    t[:, 0] = BOUND
    t[:, -1] = BOUND
    if out_corner:
        a, b = nx // 3, ny - ny // 3
        t[:a, b:] = OUT
        t[:a + 1, b - 1] = BOUND
        t[a, b - 1:] = BOUND
    for _ in range(obstacles):
        i, j = rng.integers(2, nx - 2), rng.integers(2, ny - 2)
        if t[i, j] == IN:
            t[i, j] = BOUND
    return t


def domain(t, dx, dy, seed=0):
    rng = np.random.default_rng(seed)
    nx, ny = t.shape
    vx = np.where(t == VALVE, 0.5 + rng.random(t.shape), 0.0)
    vy = np.where(t == VALVE, 0.1 * rng.standard_normal(t.shape), 0.0)
    T = np.ones(t.shape)
    return Domain2D(nx, ny, dx, dy, 1.0, t.ravel().copy(), np.zeros(t.size, np.uint8), vx.ravel(),
                    vy.ravel(), T.ravel())


def params(v_vis):
    return FluidParams(1.0, v_vis, 0.0, 0.0)


def test_solve_uv_closed_form():
    nx, ny, dx, dy, dt, nu = 9, 8, 0.1, 0.2, 0.05, 0.3
    t = walled(nx, ny, 0, out_corner=False, obstacles=0)
    o = oracle_lib.Oracle2D(domain(t, dx, dy), "f64", params(nu), method="stable")
    x = (np.arange(nx) * dx)[:, None] + 0 * np.arange(ny)[None, :]
    y = 0 * np.arange(nx)[:, None] + (np.arange(ny) * dy)[None, :]
    tu = 0.3 + 1.1 * x - 0.7 * x * x + 0.4 * y + 0.9 * y * y
    tv = -0.2 + 0.5 * x + 1.3 * x * x - 0.8 * y - 0.6 * y * y
    cu = np.cos(x + 2 * y)
    cv = np.sin(3 * x - y)
    ou, ov = o.solve_uv(dt, cu.ravel(), cv.ravel(), tu.ravel(), tv.ravel())
    ux, uy, uxx, uyy = 1.1 - 1.4 * x, 0.4 + 1.8 * y, -1.4, 1.8
    vx, vy, vxx, vyy = 0.5 + 2.6 * x, -0.8 - 1.2 * y, 2.6, -1.2
    eu = cu + dt * (-tu * ux - tv * uy + nu * (uxx + uyy))
    ev = cv + dt * (-tu * vx - tv * vy + nu * (vxx + vyy))
    m = t == IN
    np.testing.assert_allclose(ou.reshape(nx, ny)[m], eu[m], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(ov.reshape(nx, ny)[m], ev[m], rtol=1e-12, atol=1e-13)
    assert np.all(ou.reshape(nx, ny)[~m] == 0)


def _div(t, wu, wv, dx, dy):
    return np.where(t == IN, B.central(wu, dx, 0) + B.central(wv, dy, 1), 0.0)


def _w(t, seed):
    rng = np.random.default_rng(seed)
    wu = np.where(t != OUT, rng.standard_normal(t.shape), 0.0)
    wv = np.where(t != OUT, rng.standard_normal(t.shape), 0.0)
    return wu, wv


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("sweeps", [1, 2, 7])
def test_relaxation_sweeps_match_matrix_splitting(seed, sweeps):
    nx, ny, dx, dy = 11, 9, 0.03, 0.05
    t = walled(nx, ny, seed)
    if seed % 2:
        t[0, 3] = BOUND  # boundary cells on the grid's edge read q = 0 beyond it
    o = oracle_lib.Oracle2D(domain(t, dx, dy), "f64", params(0.01), method="stable")
    wu, wv = _w(t, seed)
    o.set_poisson(0.0, sweeps)
    q, pu, pv, n, capped = o.project(wu.ravel(), wv.ravel())
    want, nm = B.relax(t, dx, dy, _div(t, wu, wv, dx, dy), tol=0.0, cap=sweeps)
    assert n == sweeps == nm
    scale = np.abs(want).max()
    assert scale > 0
    np.testing.assert_allclose(q.reshape(nx, ny), want, rtol=0, atol=1e-13 * scale)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("tol", [1e-2, 1e-5])
def test_stopping_fold_and_projection(seed, tol):
    nx, ny, dx, dy = 10, 12, 0.04, 0.03
    t = walled(nx, ny, 10 + seed)
    o = oracle_lib.Oracle2D(domain(t, dx, dy), "f64", params(0.01), method="stable")
    o.set_poisson(tol, 1 << 30)
    wu, wv = _w(t, 20 + seed)
    q, pu, pv, n, capped = o.project(wu.ravel(), wv.ravel())
    want, nm = B.relax(t, dx, dy, _div(t, wu, wv, dx, dy), tol=tol)
    assert capped == 0
    assert n == nm, (n, nm)
    scale = np.abs(want).max()
    np.testing.assert_allclose(q.reshape(nx, ny), want, rtol=0, atol=1e-12 * scale)
    m = t == IN
    np.testing.assert_allclose(pu.reshape(nx, ny)[m], (wu - B.central(want, dx, 0))[m], rtol=0,
                               atol=1e-10 * scale / dx)
    np.testing.assert_allclose(pv.reshape(nx, ny)[m], (wv - B.central(want, dy, 1))[m], rtol=0,
                               atol=1e-10 * scale / dy)


def test_zero_field_stops_after_one_sweep():
    t = walled(9, 9, 3)
    o = oracle_lib.Oracle2D(domain(t, 0.1, 0.1), "f64", params(0.01), method="stable")
    z = np.zeros(t.size)
    q, pu, pv, n, capped = o.project(z, z)
    assert n == 1 and capped == 0  # every change 0/0: the fold ends NaN
    assert np.all(q == 0)
    assert B.fold_err(np.zeros(5), np.zeros(5)) != B.fold_err(np.zeros(5), np.zeros(5))


def test_fold_semantics():
    nan = float("nan")
    # NaN in the middle: the running value restarts after it
    assert B.fold_err(np.array([5.0, 0.0, 2.0]), np.array([0.0, 0.0, 1.0])) == 0.5
    # no NaN: the largest change
    assert B.fold_err(np.array([2.0, 4.0]), np.array([1.0, 1.0])) == 0.75
    # NaN last: NaN, which ends the loop (NaN >= tol is false)
    assert np.isnan(B.fold_err(np.array([2.0, 0.0]), np.array([1.0, 0.0])))
    assert not (nan >= 1e-2)


@pytest.mark.parametrize("G", [1, 2])
def test_whole_steps_match_model(G):
    nx, ny, dx, dy, dt, nu = 12, 10, 0.05, 0.05, 2e-3, 0.02
    t = walled(nx, ny, 5)
    d = domain(t, dx, dy, seed=7)
    o = oracle_lib.Oracle2D(d, "f64", params(nu), method="stable")
    gv = {"u": d.vx.reshape(nx, ny), "v": d.vy.reshape(nx, ny), "T": d.T.reshape(nx, ny)}
    m = B.StableModel(t, gv, dx, dy, 1.0, nu)
    for step in range(3):
        o.update_boundaries()
        m.update_boundaries()
        r, err, it = o.step(dt, G, 1)
        em, im = m.step(dt, G)
        assert r == 0
        assert it == im
        assert list(o.poisson_sweeps()) == m.sweeps
        np.testing.assert_allclose(err, em, rtol=1e-9)
        u, v, T = o.layer(0)
        scale = max(np.abs(m.cur["u"]).max(), np.abs(m.cur["v"]).max())
        np.testing.assert_allclose(u.reshape(nx, ny), m.cur["u"], rtol=0, atol=1e-11 * scale)
        np.testing.assert_allclose(v.reshape(nx, ny), m.cur["v"], rtol=0, atol=1e-11 * scale)
        np.testing.assert_array_equal(T.reshape(nx, ny), m.cur["T"])


def _two_chambers(nx=14, ny=9, quiet_last=True):
    """Two chambers split by a double wall (rows w, w+1): the wall rows read
    only their own side (a boundary cell takes the opposite neighbour of a
    non-NODE_IN one), so a field confined to one chamber leaves the other at
    q = 0."""
    t = walled(nx, ny, 0, out_corner=False, obstacles=0)
    w = nx // 2
    t[w, :] = BOUND
    t[w + 1, :] = BOUND
    rng = np.random.default_rng(4)
    wu = rng.standard_normal((nx, ny))
    wv = rng.standard_normal((nx, ny))
    active = np.arange(nx)[:, None] < w if quiet_last else np.arange(nx)[:, None] > w + 1
    wu = np.where(active & (t != OUT), wu, 0.0)
    wv = np.where(active & (t != OUT), wv, 0.0)
    return t, wu, wv


@pytest.mark.parametrize("quiet_last", [True, False])
def test_nan_change_of_last_cell_ends_relaxation(quiet_last):
    """With the quiet chamber last in the sweep order every sweep ends on a
    0/0 change: the fold is NaN and the relaxation stops after one sweep
    (a fold that skipped NaNs would go on).  With it first, the NaNs are
    followed by real changes and the relaxation runs to the tolerance."""
    dx = dy = 0.05
    t, wu, wv = _two_chambers(quiet_last=quiet_last)
    o = oracle_lib.Oracle2D(domain(t, dx, dy), "f64", params(0.01), method="stable")
    q, pu, pv, n, capped = o.project(wu.ravel(), wv.ravel())
    want, nm = B.relax(t, dx, dy, _div(t, wu, wv, dx, dy))
    assert n == nm
    if quiet_last:
        assert n == 1
    else:
        assert n > 5
    np.testing.assert_allclose(q.reshape(t.shape), want, rtol=0, atol=1e-12 * np.abs(want).max())
