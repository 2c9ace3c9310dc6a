"""GPU geometry ingestion (csrc/geometry.hip through cmcx_geometry_*) against
the oracle's sequential restatement of the reference's loaders: node arrays
(type, BC kinds, node velocity and temperature) bit-identical, on the
reference's example inputs (tests/golden/geometry), on two-frame files built
at times between their frames, and on polylines that cross (cells reached by
several segments take the last one in the reference's order).  Then the
example grids run end to end: built on the GPU, stepped on the GPU against the
oracle's steps of the same grid (FP64 within 1e-11, FP32 within 1e-5
velocity-scaled relative L2), output through GetLayer into a COARDS file that
an independent netCDF reader reads back."""
import os

import numpy as np
import pytest

import oracle_lib
from cmcx.geometry import Grid3D
from cmcx.solver import make_solver
from gpu_util import rel_l2_scaled, say, vel_norm

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "geometry")

CASES = {
    "box_pipe_3D": ("shape3d", "box_pipe_3D_data.txt", dict(dx=0.01, dy=0.01, dz=0.01)),
    "tetra": ("shape3d", "tetra_data.txt", dict(dx=0.01, dy=0.01, dz=0.01)),
    "box_pipe_2D": ("shape2d", "box_pipe_2D_data.txt", dict(dx=0.02, dy=0.02, dz=0.02, depth=1.0)),
    "non_uniform_pipe": ("shape2d", "non_uniform_pipe_2D_data.txt",
                         dict(dx=0.02, dy=0.02, dz=0.02, depth=1.0, depth_var=0.2)),
}


def _pair(kind, path, kw, base_t=1.0, align=False):
    o = oracle_lib.OracleGeometry(kind, path, base_t=base_t, align=align, **kw)
    g = Grid3D(kind, path, kw["dx"], kw["dy"], kw["dz"], baseT=base_t, depth=kw.get("depth", 0.0),
               depth_var=kw.get("depth_var", 0.0), align=align)
    assert g.shape == o.shape, (g.shape, o.shape)
    return g, o


def _same(g, o, time):
    got = g.arrays(time)
    want = o.build(time)
    for nm, a, b in zip(("type", "bc_vel", "bc_temp", "vx", "vy", "vz", "T"), got, want):
        nd = int(np.count_nonzero(a != b))
        assert nd == 0, f"{nm}: {nd} nodes differ"
    return got


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("align", [False, True])
def test_examples_bit_exact(name, align):
    kind, fn, kw = CASES[name]
    g, o = _pair(kind, os.path.join(GOLD, fn), kw, base_t=1.25, align=align)
    say(f"[geometry] {name} align={align} {g.shape}")
    t = _same(g, o, 0.0)[0]
    assert (t == 0).sum() > 0
    g.close()


def _two_frame_mesh(path):
    """A cube that grows and moves between two frames."""
    cube = np.array([[0, 0, 0], [0, 0, 1], [0, 1, 1], [0, 1, 0], [1, 0, 0], [1, 0, 1], [1, 1, 1], [1, 1, 0]],
                    np.float64)
    tris = [(0, 1, 2), (0, 2, 3), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4), (3, 2, 6), (3, 6, 7),
            (0, 4, 7), (0, 7, 3), (1, 5, 6), (1, 6, 2)]
    with open(path, "w") as f:
        f.write("2\n")
        for scale, shift in ((700.0, 100.0), (800.0, 180.0)):
            f.write("8\n")
            for v in cube * scale + shift:
                f.write(f"{v[0]:.1f} {v[1]:.1f} {v[2]:.1f} 0.0 0.0 0.0\n")
            f.write("12\n")
            for t in tris:
                f.write(f"{t[0]} {t[1]} {t[2]}\n")


def _crossing_polylines(path, frames=2):
    """A square channel (valves left and right, walls top and bottom) with a
    diagonal wall and a polyline crossing it, the top wall moving between frames."""
    with open(path, "w") as f:
        f.write(f"{frames}\n")
        for k in range(frames):
            top = 1000.0 + 150.0 * k
            f.write("0.02\n5\n")
            f.write(f"2\n0.0 0.0\n1000.0 0.0\nPassive\n")
            f.write(f"2\n1000.0 {top}\n0.0 {top}\nPassive\n")
            f.write(f"2\n0.0 0.0\n0.0 {top}\nMotion\n800.0 0.0\n")
            f.write(f"2\n1000.0 0.0\n1000.0 {top}\nMotion\n0.0 0.0\n")
            f.write("3\n200.0 200.0\n700.0 650.0\n300,0 800,0\nPassive\n")


def test_two_frame_mesh(tmp_path):
    p = tmp_path / "mesh2.txt"
    _two_frame_mesh(p)
    g, o = _pair("shape3d", str(p), dict(dx=0.02, dy=0.02, dz=0.02))
    say(f"[geometry] two-frame mesh {g.shape}")
    for time in (0.0, 0.4 / 75, 0.77 / 75, 1.5 / 75):
        _same(g, o, time)
    g.close()


def test_crossing_polylines(tmp_path):
    p = tmp_path / "poly2.txt"
    _crossing_polylines(p)
    g, o = _pair("shape2d", str(p), dict(dx=0.02, dy=0.02, dz=0.025, depth=0.6, depth_var=0.3))
    say(f"[geometry] crossing polylines {g.shape}")
    for time in (0.0, 0.005, 0.013, 0.031):
        _same(g, o, time)
    g.close()


@pytest.mark.parametrize("name,dtype,steps", [("box_pipe_2D", "f64", 6), ("box_pipe_2D", "f32", 6),
                                              ("tetra", "f64", 3), ("box_pipe_3D", "f32", 3)])
def test_example_runs_end_to_end(name, dtype, steps, tmp_path):
    """The reference's example configs (Re 200, Pr 0.72, lambda 1.4, 10 steps per
    frame of 1/75 s, num_global 4, num_local 2): grid on the GPU, steps on the
    GPU against the oracle, the state's output view to a COARDS file."""
    from cmcx.io import OutputNetCDF3D_header, OutputNetCDF3D_layer

    netcdf = pytest.importorskip("scipy.io").netcdf_file
    kind, fn, kw = CASES[name]
    g = Grid3D(kind, os.path.join(GOLD, fn), kw["dx"], kw["dy"], kw["dz"], baseT=1.0,
               depth=kw.get("depth", 0.0), depth_var=kw.get("depth_var", 0.0))
    dt = g.cycle_length / (g.num_frames * 10)
    d = g.Prepare(0.0, re=200.0, dt=dt, num_global=4, num_local=2)
    say(f"[end to end] {name} {dtype} {d.shape}, dt {dt:.3e}, {steps} steps")
    ref = oracle_lib.Oracle3D(d, "f64")
    ref.create_segments()
    s = make_solver(d, dtype, path=1)
    out = (max(2, d.nx // 2), max(2, d.ny // 2), max(2, d.nz // 2))
    nc = tmp_path / "res.nc"
    OutputNetCDF3D_header(nc, g.bbox_min, g.bbox_max, dt * 2, dt * steps, *out, ["u", "v", "w", "T"])
    layers = []
    for i in range(steps):
        ref.update_boundaries()
        s.UpdateBoundaries()
        r, _ = ref.step(dt, 4, 2, False)
        assert r == 0
        s.TimeStep(dt, 4, 2, i == steps - 1)
        if i % 2 == 0:
            # GetLayer writes 99999 into the NODE_OUT nodes of the buffer that
            # becomes the next state (Solver3D.cpp:21-25); the oracle's
            # get_layer does the same, so both runs carry it.
            vel, T = s.GetLayer(*out)
            rvel, rT = ref.get_layer(*out)
            tl = 1e-11 if dtype == "f64" else 1e-5
            np.testing.assert_allclose(T, rT, rtol=tl, atol=tl)
            np.testing.assert_allclose(vel, rvel, rtol=tl, atol=tl * max(1.0, float(np.abs(rvel).max())))
            OutputNetCDF3D_layer(nc, vel, T, i // 2, *out, "u v w T")
            layers.append((vel, T))
    u, v, w, p, T = s.fields()
    ru, rv, rw, rT = ref.fields()
    # NODE_OUT nodes carry no solution: they hold what GetLayer's fill left in
    # each state buffer, the same in both runs
    fluid = d.type.reshape(-1) != 1
    wd = np.float32 if dtype == "f32" else np.float64
    for nm, a, b in (("u", u, ru), ("v", v, rv), ("w", w, rw), ("T", T, rT)):
        a, b = a.reshape(-1), b.reshape(-1)
        assert np.array_equal(a[~fluid], b[~fluid].astype(wd)), f"{nm}: NODE_OUT values differ"
    if steps > 4:
        assert (u.reshape(-1)[~fluid] == 99999.0).all()
    ru, rv, rw, rT = (x.reshape(-1)[fluid] for x in (ru, rv, rw, rT))
    u, v, w, T = (x.reshape(-1)[fluid] for x in (u, v, w, T))
    vs = vel_norm(ru, rv, rw)
    assert vs > 0
    tol = 1e-11 if dtype == "f64" else 1e-5
    for nm, a, b, sc in (("u", u, ru, vs), ("v", v, rv, vs), ("w", w, rw, vs), ("T", T, rT, 0.0)):
        e = rel_l2_scaled(a, b, sc)
        say(f"  {nm}: rel L2 vs FP64 oracle {e:.3e}")
        assert e < tol, nm
    with netcdf(str(nc), "r", mmap=False, maskandscale=False) as f:
        for k, (vel, Tl) in enumerate(layers):
            np.testing.assert_array_equal(f.variables["u"].data[k].ravel(), vel[:, 0].astype(np.float64))
            np.testing.assert_array_equal(f.variables["T"].data[k].ravel(), Tl)
        if (d.type == 1).any():  # NODE_OUT nodes read 99999 in the output view
            assert (f.variables["T"].data[0] == 99999.0).any()
    s.close()
    g.close()
