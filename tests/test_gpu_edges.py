"""GPU edge cases through the C ABI against the oracle: a domain with no NODE_IN
node (empty segment lists and work lists), the longest lines each fused kernel
takes (Z lines of 512 nodes in FP32 -- 16 rows per lane, or 8 on 64 lanes --
and of 256 in FP64; X and Y lines of 256), and the rejection of longer lines
with CMCX_ERR_UNSUPPORTED."""
import numpy as np
import pytest

import oracle_lib
from cmcx import domains
from cmcx._lib import CmcxError
from cmcx.solver import make_solver
from gpu_util import rel_l2, rel_l2_scaled, say, vel_norm

pytestmark = pytest.mark.gpu

PLAIN, FUSED = 0, 1


def _no_fluid(shape, seed=0):
    """A box whose every node is a wall (NODE_BOUND), with random node values."""
    rng = np.random.default_rng(seed)
    d = domains.box(*shape)
    t = d.type.reshape(shape)
    t[t == domains.NODE_IN] = domains.NODE_BOUND
    d.vx = rng.normal(size=d.n)
    d.vy = rng.normal(size=d.n)
    d.vz = rng.normal(size=d.n)
    d.T = 1.0 + 0.1 * rng.normal(size=d.n)
    return d


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("path", [PLAIN, FUSED])
def test_no_fluid_node(path, dtype):
    """No segment anywhere: the step only carries the boundary fill through, and
    EvalDivError divides 0 by 0 like the reference (TimeLayer3D.h:595-641),
    which its guard (err > 0.01) lets pass."""
    d = _no_fluid((9, 7, 12))
    say(f"[no fluid] path={path} {dtype}")
    ref = oracle_lib.Oracle3D(d, dtype)
    ref.create_segments()
    s = make_solver(d, dtype, path=path)
    for _ in range(2):
        ref.update_boundaries()
        s.UpdateBoundaries()
        r, e_ref = ref.step(d.dt, 2, 1, True)
        assert r == 0
        e = s.TimeStep(d.dt, 2, 1, True)
        assert np.isnan(e) and np.isnan(e_ref)
    got = s.fields()
    want = ref.fields()
    for k, idx in enumerate((0, 1, 2, 4)):
        assert np.array_equal(got[idx], want[k]), k
    assert np.array_equal(got[0], d.vx.astype(got[0].dtype))
    s.close()


def _long(shape, seed):
    """Random solids and a random initial state on the NODE_IN nodes."""
    rng = np.random.default_rng(seed)
    d = domains.random_domain(shape, rng, 0.12)
    inner = d.type == domains.NODE_IN
    for a in (d.vx, d.vy, d.vz):
        a[inner] = rng.normal(0.0, 0.3, int(inner.sum()))
    d.T[inner] = 1.0 + 0.1 * rng.normal(size=int(inner.sum()))
    return d


def _compare(d, dtype, path, variants, steps=2, tol=None):
    ref = oracle_lib.Oracle3D(d, "f64")
    ref.create_segments()
    for _ in range(steps):
        ref.update_boundaries()
        r, _ = ref.step(d.dt, 2, 1, False)
        assert r == 0
    want = ref.fields()
    vs = vel_norm(*want[:3])
    tol = tol or (1e-11 if dtype == "f64" else 1e-5)
    for v in variants:
        s = make_solver(d, dtype, path=path)
        for a, va in enumerate(v):
            s.set_tile_variant(a, va)
        for _ in range(steps):
            s.UpdateBoundaries()
            s.TimeStep(d.dt, 2, 1, False)
        got = s.fields()
        for k in range(3):
            e = rel_l2_scaled(got[k], want[k], vs)
            say(f"  {d.shape} {dtype} variants {v} field {k}: {e:.3e}")
            assert e < tol, (v, k)
        e = rel_l2(got[4], want[3])
        assert e < tol, (v, "T")
        s.close()


@pytest.mark.parametrize("dtype,nz", [("f32", 512), ("f32", 256), ("f64", 256)])
def test_longest_z_lines(dtype, nz):
    """Every Z variant at the longest z lines the fused path takes, and in FP32
    at nz = 256 (the 256^3 workload's lines: 4 rows per lane over 64 lanes or 8
    over 32, whole 16-byte runs)."""
    d = _long((10, 9, nz), 3)
    say(f"[long z] {dtype} nz={nz}")
    _compare(d, dtype, FUSED, [(0, 0, z) for z in range(8)])


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_longest_strided_lines(dtype):
    """X and Y lines of 256 nodes: the 128-, 64- and 32-byte tile variants, with
    the Picard layer re-read (0-2) and parked in LDS (6-8; the 128-byte one
    fills the 160 KB of LDS exactly)."""
    for shape in ((256, 9, 12), (9, 256, 12)):
        d = _long(shape, 5)
        say(f"[long strided] {dtype} {shape}")
        _compare(d, dtype, FUSED, [(v, v, 0) for v in (0, 1, 2, 6, 7, 8)])


@pytest.mark.parametrize("dtype,shape", [("f32", (8, 8, 1024)), ("f64", (8, 8, 512)),
                                         ("f32", (512, 8, 8)), ("f64", (8, 512, 8))])
def test_longer_lines_rejected(dtype, shape):
    """Lines beyond the fused kernels' reach: the fused path is refused with
    CMCX_ERR_UNSUPPORTED (the plain path takes them)."""
    d = _long(shape, 7)
    with pytest.raises(CmcxError):
        make_solver(d, dtype, path=FUSED)
    s = make_solver(d, dtype, path=PLAIN)
    s.UpdateBoundaries()
    s.TimeStep(d.dt, 1, 1, False)
    s.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("graph", [False, True])
def test_boundary_values_out_of_sync(dtype, graph):
    """Boundary rows take the node values (Grid3D Node data), whatever the state
    holds on BOUND/VALVE nodes: after set_fields and after set_boundary without
    UpdateBoundaries the state differs from the node values there, and the
    fused sweeps read the node-value arrays; after UpdateBoundaries they read
    the same values from the Picard layer.  Both against the oracle."""
    rng = np.random.default_rng(21)
    d = domains.random_domain((14, 12, 16), rng, 0.12, free_faces=True)
    n = d.n
    say(f"[bsync] {dtype} graph={graph}")
    ref = oracle_lib.Oracle3D(d, "f64")
    ref.create_segments()
    s = make_solver(d, dtype, path=FUSED)
    s.set_graph(graph)
    tol = 1e-11 if dtype == "f64" else 1e-5

    def check(tag):
        got = s.fields()
        want = ref.fields()
        vs = vel_norm(*want[:3])
        for k in range(3):
            e = rel_l2_scaled(got[k], want[k], vs)
            say(f"  {tag} field {k}: {e:.3e}")
            assert e < tol, (tag, k)
        assert rel_l2(got[4], want[3]) < tol, (tag, "T")

    # 1. a state that differs from the node values everywhere, no boundary fill
    st = [rng.normal(0.0, 0.3, n) for _ in range(3)] + [1.0 + 0.1 * rng.normal(size=n)]
    s.set_fields(*st)
    ref.set_fields(*st)
    for _ in range(2):
        s.TimeStep(d.dt, 2, 1, False)
        r, _ = ref.step(d.dt, 2, 1, False)
        assert r == 0
    check("set_fields, no fill")
    # 2. new node values, no boundary fill
    bv = [d.vx + 0.2 * rng.normal(size=n), d.vy + 0.2 * rng.normal(size=n),
          d.vz + 0.2 * rng.normal(size=n), d.T + 0.1 * rng.normal(size=n)]
    s.set_boundary(d.bc_vel, d.bc_temp, *bv)
    ref.set_boundary(d.bc_vel, d.bc_temp, *bv)
    s.TimeStep(d.dt, 2, 1, False)
    r, _ = ref.step(d.dt, 2, 1, False)
    assert r == 0
    check("set_boundary, no fill")
    # 3. boundary fill, then steps in sync
    for _ in range(2):
        s.UpdateBoundaries()
        ref.update_boundaries()
        s.TimeStep(d.dt, 2, 1, False)
        r, _ = ref.step(d.dt, 2, 1, False)
        assert r == 0
    check("filled")
    s.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("path,graph", [(FUSED, False), (FUSED, True), (PLAIN, False)])
def test_short_boundary_lists_bit_identical(dtype, path, graph):
    """While the state is in sync with the node values, the boundary fill and the
    step set-up visit only the BOUND/VALVE nodes with a free BC.  Solver b gets
    the same node values again before every step, which sends it through the
    whole lists with Dirichlet values read from the node-value arrays; the two
    runs must agree bit for bit, on every node (no-slip rows write back exactly
    the value they read).  Some nodes are free for T only, and one step has no
    boundary fill."""
    rng = np.random.default_rng(33)
    d = domains.random_domain((15, 13, 17), rng, 0.12, free_faces=True)
    inner = d.type == domains.NODE_IN
    for a in (d.vx, d.vy, d.vz):
        a[inner] = rng.normal(0.0, 0.3, int(inner.sum()))
    d.T[inner] = 1.0 + 0.1 * rng.normal(size=int(inner.sum()))
    wall = np.flatnonzero((d.type == domains.NODE_BOUND) & (d.bc_temp == domains.BC_NOSLIP))
    d.bc_temp[rng.choice(wall, wall.size // 5, replace=False)] = domains.BC_FREE
    d.T[d.type != domains.NODE_IN] += 0.05 * rng.normal(size=int((~inner).sum()))
    say(f"[short lists] {dtype} path={path} graph={graph}")
    a = make_solver(d, dtype, path=path)
    b = make_solver(d, dtype, path=path)
    a.set_graph(graph)
    b.set_graph(graph)
    for i in range(5):
        b.set_boundary(d.bc_vel, d.bc_temp, d.vx, d.vy, d.vz, d.T)
        if i != 3:
            a.UpdateBoundaries()
            b.UpdateBoundaries()
        L = 2 if i == 1 else 1
        a.TimeStep(d.dt, 2, L, False)
        b.TimeStep(d.dt, 2, L, False)
    fa, fb = a.fields(), b.fields()
    for k in (0, 1, 2, 4):
        bad = np.flatnonzero(fa[k] != fb[k])
        assert bad.size == 0, (k, bad[:5], d.type[bad[:5]])
    ref = oracle_lib.Oracle3D(d, "f64")
    ref.create_segments()
    for i in range(5):
        if i != 3:
            ref.update_boundaries()
        r, _ = ref.step(d.dt, 2, 2 if i == 1 else 1, False)
        assert r == 0
    want = ref.fields()
    vs = vel_norm(*want[:3])
    tol = 1e-11 if dtype == "f64" else 1e-5
    for k in range(3):
        assert rel_l2_scaled(fa[k], want[k], vs) < tol, k
    assert rel_l2(fa[4], want[3]) < tol
    a.close()
    b.close()


@pytest.mark.parametrize("path", [PLAIN, FUSED])
def test_blown_up_state_fires_guard(path):
    """A NaN in the state (a run that blew up) fires the divergence guard: the
    step returns CMCX_DIVERGED and leaves the state unadvanced.  The reference
    tests err > 0.01, which a NaN passes; a grid without fluid nodes (0/0, see
    test_no_fluid_node) still passes.  In a batch, a NaN in any grid fires it."""
    from cmcx._lib import DivergenceError
    from cmcx.solver import make_batch

    d = domains.lid_sphere(24, radius=5.0, dt=1e-3)
    say(f"[blown-up state] path={path}")
    s = make_solver(d, "f32", path=path)
    u, v, w, p, T = s.fields()
    fluid = np.flatnonzero(d.type == domains.NODE_IN)
    u[fluid[len(fluid) // 2]] = np.nan
    s.set_fields(u, v, w, T)
    with pytest.raises(DivergenceError):
        s.TimeStep(d.dt, 2, 1, True)
    got = s.fields()
    assert np.isnan(got[0][fluid[len(fluid) // 2]])  # unadvanced: still the state set
    assert np.array_equal(got[1], v)
    s.close()
    doms = domains.batch_boxes(count=3, seed=2, edges=(12, 16))
    b = make_batch(doms, "f32", path=path)
    b.select(2)
    u, v, w, p, T = b.fields()
    fl = np.flatnonzero(doms[2].type == domains.NODE_IN)
    T[fl[0]] = np.inf
    b.set_fields(u, v, w, T)
    with pytest.raises(DivergenceError):
        b.TimeStep(doms[0].dt, 2, 1, True)
    errs = b.grid_errors()
    assert np.all(np.isfinite(errs[:2])) and not np.isfinite(errs[2])
    b.close()
