"""Oracle pins of the geometry ingestion (oracle/geometry.cpp, the reference's
Shape3D / Shape2D loaders and builds) on the reference's example inputs
(tests/golden/geometry), without a GPU.

* The 1000 mm cube (box_pipe_3D): with the 2 % box padding and dx = 10 mm
  its faces land on grid planes 2 and 102 of a 105^3 grid: walls exactly on
  the cube surface [2, 102]^3, fluid exactly in [3, 101]^3, outside the rest;
  fluid and outside nodes take T = baseT, walls T = 0.
* The tetrahedron: every fluid node lies inside the tetrahedron (a point test
  in grid coordinates with half a cell of slack), the fluid volume is close to
  the tetrahedron's, walls lie within a cell of a face plane.
* The extruded square (box_pipe_2D): the 2D grid has walls on rows j = 1 and
  j = 51, valves on columns i = 1 (inflow, u = 1 m/s) and i = 51 (u = 0: free
  BCs), the corners walls (walls are drawn after valves), fluid inside, and one
  wall cell at (0, 51) that single-precision stepping of the top wall reaches; the
  columns follow Prepare2D's layers: k = 0 outside, k = 1 bottom wall, fluid
  k = 2..48, top wall k = 49 (free T BC), k = 50 outside.
* The non-uniform pipe (depth_var 0.2): the bottom wall of column (i, j) ends
  at 1 + int(0.2 z (47)) with z = 1 - (x^2 + y^2)/2, x = -1 + 2i/53.
* Every grid is well formed (no fluid on a box face, no outside node next to
  fluid) and a one-frame file builds the same grid at every time.
"""
import os

import numpy as np
import pytest

import oracle_lib
from cmcx import domains

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "geometry")
IN, OUT, BOUND, VALVE = 0, 1, 2, 3


def _geo(name, **kw):
    kind = "shape3d" if name in ("box_pipe_3D_data.txt", "tetra_data.txt") else "shape2d"
    return oracle_lib.OracleGeometry(kind, os.path.join(GOLD, name), **kw)


def _well_formed(t):
    fin = t == IN
    assert not (fin[0].any() or fin[-1].any() or fin[:, 0].any() or fin[:, -1].any()
                or fin[:, :, 0].any() or fin[:, :, -1].any())
    out = t == OUT
    for ax in range(3):
        for sh in (1, -1):
            assert not (out & np.roll(fin, sh, axis=ax)).any()


def test_box_3d_analytic():
    g = _geo("box_pipe_3D_data.txt", dx=0.01, dy=0.01, dz=0.01, base_t=1.5)
    assert g.shape == (105, 105, 105)
    t, bv, bt, vx, vy, vz, T = g.build(0.0)
    t = t.reshape(g.shape)
    want = np.full(g.shape, OUT, np.uint8)
    want[2:103, 2:103, 2:103] = BOUND
    want[3:102, 3:102, 3:102] = IN
    np.testing.assert_array_equal(t, want)
    T = T.reshape(g.shape)
    assert np.all(T[want != BOUND] == 1.5) and np.all(T[want == BOUND] == 0.0)
    assert np.all(bv == 0) and np.all(bt == 0) and not (vx.any() or vy.any() or vz.any())
    _well_formed(t)


def test_tetra_inside():
    g = _geo("tetra_data.txt", dx=0.01, dy=0.01, dz=0.01)
    t = g.build(0.0)[0].reshape(g.shape)
    _well_formed(t)
    # the vertices in grid coordinates (mm * 0.001, 2 % padding, / 0.01)
    P = np.array([[0, 500, 0], [0, 0, 500], [1000, 0, 0], [0, 500, 200]], np.float64) * 1e-3
    lo, hi = P.min(0), P.max(0)
    pad = (hi - lo) * 0.02
    V = (P - (lo - pad)) / 0.01
    faces = [(0, 2, 3), (1, 2, 3), (0, 1, 3), (0, 1, 2)]
    c = V.mean(0)

    def inside(X, slack):
        ok = np.ones(len(X), bool)
        for f in faces:
            a, b, d = V[list(f)]
            n = np.cross(b - a, d - a)
            n /= np.linalg.norm(n)
            s = np.sign(np.dot(c - a, n))
            ok &= s * ((X - a) @ n) >= -slack
        return ok

    fl = np.argwhere(t == IN).astype(np.float64)
    assert len(fl) > 1000
    assert inside(fl, 0.5).all()
    vol = abs(np.dot(np.cross(V[1] - V[0], V[2] - V[0]), V[3] - V[0])) / 6
    assert 0.5 * vol < len(fl) < vol
    wall = np.argwhere(t == BOUND).astype(np.float64)
    dist = np.full(len(wall), np.inf)
    for f in faces:
        a, b, d = V[list(f)]
        n = np.cross(b - a, d - a)
        n /= np.linalg.norm(n)
        dist = np.minimum(dist, np.abs((wall - a) @ n))
    assert dist.max() < 1.8


@pytest.mark.parametrize("name,dv", [("box_pipe_2D_data.txt", 0.0), ("non_uniform_pipe_2D_data.txt", 0.2)])
def test_extruded_square(name, dv):
    g = _geo(name, dx=0.02, dy=0.02, dz=0.02, depth=1.0, depth_var=dv, base_t=1.0)
    assert g.shape == (53, 53, 51)
    t, bv, bt, vx, vy, vz, T = (a.reshape(g.shape) for a in g.build(0.0))
    c2 = g.grid2d()
    want2 = np.full((53, 53), OUT, np.uint8)
    want2[1:52, 1:52] = IN
    want2[1, 1:52] = VALVE
    want2[51, 1:52] = VALVE
    want2[1:52, 1] = BOUND
    want2[1:52, 51] = BOUND
    # the top wall runs from x = 1000 to x = 0 mm: 51 steps of -50/51 cells
    # accumulated in single precision end just below 1.0, whose truncation
    # draws one cell more, outside the box (the reference's arithmetic)
    p = np.float32(51.0)
    for _ in range(51):
        p = np.float32(p + np.float32(np.float32(-50.0) / np.float32(51)))
    assert int(p) == 0
    want2[0, 51] = BOUND
    np.testing.assert_array_equal(c2, want2)
    _well_formed(t)
    az = 51
    for i in range(53):
        for j in range(53):
            col = t[i, j]
            if want2[i, j] == OUT:
                assert np.all(col == OUT)
                continue
            assert col[0] == OUT and col[az - 1] == OUT
            x, y = -1 + 2 * i / 53, -1 + 2 * j / 53
            bottom = 1 + int(dv * (1.0 - (x * x + y * y) * 0.5) * (az - 4))
            assert np.all(col[1:bottom + 1] == BOUND) and col[az - 2] == BOUND
            mid = col[bottom + 1:az - 2]
            assert np.all(mid == {IN: IN, BOUND: BOUND, VALVE: VALVE}[want2[i, j]])
            assert bt[i, j, az - 2] == 1 and bv[i, j, az - 2] == 0  # no-slip walls, free T
    # the inflow valve carries u = 1 m/s with no-slip BCs, the outflow valve u = 0 with free BCs
    inflow = (t == VALVE) & (np.arange(53)[:, None, None] == 1)
    outflow = (t == VALVE) & (np.arange(53)[:, None, None] == 51)
    assert np.all(vx[inflow] == np.float32(1.0)) and np.all(bv[inflow] == 0) and np.all(bt[inflow] == 0)
    assert np.all(vx[outflow] == 0) and np.all(bv[outflow] == 1) and np.all(bt[outflow] == 1)
    assert np.all(T[t == IN] == 1.0)


@pytest.mark.parametrize("name", ["box_pipe_3D_data.txt", "box_pipe_2D_data.txt"])
def test_one_frame_time_invariant(name):
    kw = dict(dx=0.02, dy=0.02, dz=0.02, depth=1.0) if "2D" in name else dict(dx=0.01, dy=0.01, dz=0.01)
    g = _geo(name, **kw)
    a = g.build(0.0)
    b = g.build(0.004)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_grid_feeds_the_solver_domain():
    """The oracle's grid is a valid solver domain (the mask the solver requires)."""
    g = _geo("box_pipe_2D_data.txt", dx=0.02, dy=0.02, dz=0.02, depth=1.0)
    t = g.build(0.0)[0]
    segs = oracle_lib.segments(t, g.shape, 0)
    assert len(segs) > 0 and np.all(segs[:, 6] >= 3)
