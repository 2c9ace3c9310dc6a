"""Pins of the oracle (CPU, no GPU).

Each check is chosen so that a plausible slip in the restatement -- a dropped
term, a wrong sign or index, a transposed operand -- fails it:
  * Thomas and FluidParams: bit-exact against the reference's own headers
    compiled from /root/reference (oracle/_ref), and Thomas against a dense
    numpy solve;
  * segment builder: an independent numpy run-length restatement, hand-made
    known answers, and invariants;
  * row assembly: closed forms of single steps (diffusion eigenmode decay,
    pressure-gradient sign per axis, advection sign per axis, every term of the
    dissipation function);
  * the composed step: a dense brute force written from the model
    (tests/brute_adi.py) on tiny domains.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib
from cmcx import domains
from cmcx.params import FluidParams

_f = ctypes.POINTER(ctypes.c_float)
_d = ctypes.POINTER(ctypes.c_double)


def _sys(n, rng, dtype, bc0="noslip", bc1="noslip"):
    a = rng.uniform(-1, 1, n).astype(dtype)
    c = rng.uniform(-1, 1, n).astype(dtype)
    b = (np.abs(a) + np.abs(c) + rng.uniform(0.5, 2, n)).astype(dtype)
    d = rng.uniform(-1, 1, n).astype(dtype)
    if bc0 == "free":
        b[0], c[0], d[0] = 2, -1, 0
    else:
        b[0], c[0] = 1, 0
    if bc1 == "free":
        a[-1], b[-1], d[-1] = -1, 2, 0
    else:
        a[-1], b[-1] = 0, 1
    return a, b, c, d


# ---------------------------------------------------------------- Thomas ----
@pytest.mark.parametrize("n", [3, 4, 17, 256])
@pytest.mark.parametrize("bcs", [("noslip", "noslip"), ("free", "noslip"), ("noslip", "free"), ("free", "free")])
def test_thomas_matches_dense_solve(n, bcs):
    rng = np.random.default_rng(n)
    a, b, c, d = _sys(n, rng, np.float64, *bcs)
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    want = np.linalg.solve(A, d)
    x = np.zeros(n)
    L = oracle_lib.load()
    aa, bb, cc, dd = (v.copy() for v in (a, b, c, d))
    L.orc_thomas_f64(*[v.ctypes.data_as(_d) for v in (aa, bb, cc, dd, x)], n)
    assert np.allclose(x, want, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_thomas_bit_exact_vs_reference_header(prec):
    R = oracle_lib.load_ref()
    if R is None:
        pytest.skip("reference tree not present: oracle/_ref not built")
    L = oracle_lib.load()
    dt = np.float32 if prec == "f32" else np.float64
    ptr = _f if prec == "f32" else _d
    rng = np.random.default_rng(3)
    for n in (3, 5, 64, 255):
        sysm = _sys(n, rng, dt, "free", "noslip")
        x1 = np.zeros(n, dt)
        x2 = np.zeros(n, dt)
        v1 = [s.copy() for s in sysm]
        v2 = [s.copy() for s in sysm]
        getattr(R, f"ref_thomas_{prec}")(*[v.ctypes.data_as(ptr) for v in v1 + [x1]], n)
        getattr(L, f"orc_thomas_{prec}")(*[v.ctypes.data_as(ptr) for v in v2 + [x2]], n)
        assert np.array_equal(x1, x2)
        assert np.array_equal(v1[2], v2[2]) and np.array_equal(v1[3], v2[3])  # c', d' too


def test_fluid_params_bit_exact_vs_reference_header():
    R = oracle_lib.load_ref()
    if R is None:
        pytest.skip("reference tree not present: oracle/_ref not built")
    out = np.zeros(4, np.float32)
    for Re, Pr, lam in ((400.0, 0.72, 1.4), (1000.0, 0.72, 1.4), (123.4, 0.9, 1.67)):
        R.ref_fluid_params_normalized(Re, Pr, lam, out.ctypes.data_as(_f))
        p = FluidParams.normalized(Re, Pr, lam)
        assert np.array_equal(out, np.array([p.v_T, p.v_vis, p.t_vis, p.t_phi], np.float32))
    R.ref_fluid_params_dimensional(1.8e-5, 1.2, 287.0, 0.025, 718.0, out.ctypes.data_as(_f))
    p = FluidParams.dimensional(1.8e-5, 1.2, 287.0, 0.025, 718.0)
    assert np.array_equal(out, np.array([p.v_T, p.v_vis, p.t_vis, p.t_phi], np.float32))


# ------------------------------------------------------ segment builder ----
def brute_segments(t, shape, axis):
    """Independent restatement: per line (reference order), cells 1..n-1 are
    scanned; a run of NODE_IN cells closed by a non-IN cell gives (start-1, end)."""
    nx, ny, nz = shape
    T = t.reshape(shape)
    out = []
    if axis == 0:
        lines = [((0, i, j), T[:, i, j]) for i in range(ny) for j in range(nz)]
    elif axis == 1:
        lines = [((i, 0, j), T[i, :, j]) for i in range(nx) for j in range(nz)]
    else:
        lines = [((i, j, 0), T[i, j, :]) for i in range(nx) for j in range(ny)]
    for start, line in lines:
        isin = np.concatenate([[False], line[1:] == 0])  # cell 0 never tested
        edges = np.diff(isin.astype(np.int8))
        starts = np.nonzero(edges == 1)[0] + 1   # first IN index of each run
        ends = np.nonzero(edges == -1)[0] + 1    # first non-IN index after a run
        for s0 in starts:
            e = ends[ends > s0]
            if len(e) == 0:
                continue  # run reaching the last node: dropped
            e = e[0]
            p = list(start)
            q = list(start)
            p[axis] = s0 - 1
            q[axis] = e
            out.append(p + q + [e - (s0 - 1) + 1])
    return np.array(out, np.int32).reshape(-1, 7)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_segments_vs_brute_force(seed):
    rng = np.random.default_rng(seed)
    shape = (9, 11, 7)
    # deliberately ill-formed too: IN on faces, runs touching node 0 and the last node
    t = rng.choice([0, 0, 0, 2, 1], size=shape).astype(np.uint8).ravel()
    for axis in range(3):
        assert np.array_equal(oracle_lib.segments(t, shape, axis), brute_segments(t, shape, axis))


def test_segment_known_answers():
    # one line along X of length 8 (y, z dims 3 with solid faces)
    def line_mask(pattern):
        t = np.full((len(pattern), 3, 3), 2, np.uint8)
        t[:, 1, 1] = pattern
        return t.ravel(), t.shape

    def xsegs(pattern):
        t, shape = line_mask(pattern)
        s = oracle_lib.segments(t, shape, 0)
        return [tuple(r[[0, 3, 6]]) for r in s]

    B, I = 2, 0
    assert xsegs([B, B, B, B]) == []                          # empty line
    assert xsegs([B, I, B, B]) == [(0, 2, 3)]                 # single IN cell: size 3
    assert xsegs([B, I, I, B, I, B]) == [(0, 3, 4), (3, 5, 3)]  # one-node wall shared
    assert xsegs([I, I, I, B]) == [(0, 3, 4)]                 # node 0 never tested: starts at 0
    assert xsegs([B, B, I, I]) == []                          # run reaching the end: dropped
    assert xsegs([B, I, B, I, I]) == [(0, 2, 3)]              # second run dropped


def test_segment_invariants_well_formed():
    d = domains.ellipsoids(40, count=6, axis_lo=3, axis_hi=8, seed=3)
    t = d.type.reshape(d.shape)
    for axis in range(3):
        s = oracle_lib.segments(d.type, d.shape, axis)
        cover = np.zeros(d.shape, np.int32)
        for r in s:
            p, q, n = r[:3], r[3:6], r[6]
            assert n >= 3 and n == (q - p).sum() + 1
            assert t[tuple(p)] != 0 and t[tuple(q)] != 0       # endpoints non-IN
            idx = [slice(v, v + 1) for v in p]
            idx[axis] = slice(p[axis] + 1, q[axis])
            assert np.all(t[tuple(idx)] == 0)                  # interior all IN
            cover[tuple(idx)] += 1
        assert np.array_equal(cover, (t == 0).astype(np.int32))  # each IN node once


def test_segments_nblockz_is_a_permutation_grouped_by_z_blocks():
    rng = np.random.default_rng(5)
    d = domains.random_domain((12, 10, 13), rng, 0.2)
    for axis in (0, 1):
        s1 = oracle_lib.segments(d.type, d.shape, axis, 1)
        s3 = oracle_lib.segments(d.type, d.shape, axis, 3)
        assert sorted(map(tuple, s1)) == sorted(map(tuple, s3))
        # z-blocks [0,4),[4,8),[8,13) in order
        blk = np.searchsorted([4, 8], s3[:, 2], side="right")
        assert np.all(np.diff(blk) >= 0)


# ---------------------------------------------------------- closed forms ----
def _box(n, params=None):
    d = domains.box(n, n, n)
    return d


def _centre(a, shape, pad):
    a = np.asarray(a).reshape(shape)
    return a[pad:-pad, pad:-pad, pad:-pad]


def _run1(d, p, dt, G=1):
    o = oracle_lib.Oracle3D(d, "f64", params=p)
    o.create_segments()
    o.update_boundaries()
    r, _ = o.step(dt, G, 1, False)
    assert r == 0
    return o.fields()


def test_diffusion_eigenmode_decay():
    """T = T0 + A sin sin sin with zero velocity and v_T = 0: every sweep is a
    pure implicit diffusion with dt/3 (b = 3/dt + 2 nu/h^2); the discrete mode
    decays by prod_d 1/(1 + dt/3 * t_vis * 4/h^2 sin^2(pi h / 2L))."""
    n = 12
    d = _box(n)
    h = d.dx
    i = np.arange(n)
    s = np.sin(np.pi * i / (n - 1))
    mode = s[:, None, None] * s[None, :, None] * s[None, None, :]
    d.T = (1.0 + 1e-3 * mode).ravel()
    p = FluidParams(0.0, 1e-3, 0.05, 0.0)
    dt = 1e-2
    lam = 4.0 / h**2 * np.sin(np.pi * h / (2 * 1.0)) ** 2
    factor = (1.0 / (1.0 + dt / 3 * p.t_vis * lam)) ** 3
    for G in (1, 3):
        u, v, w, T = _run1(d, p, dt, G)
        assert np.all(u == 0) and np.all(v == 0) and np.all(w == 0)
        got = (T.reshape(n, n, n) - 1.0)
        assert np.allclose(got, 1e-3 * factor * mode, rtol=1e-10, atol=1e-16)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_pressure_term_sign_per_axis(axis):
    """T linear along one axis, zero velocity: only that axis' momentum row gets
    -v_T dT/dx; away from the walls the component becomes -(dt/3) v_T gamma."""
    n = 24
    d = _box(n)
    g = 0.7
    x = np.arange(n) * d.dx
    prof = [x[:, None, None], x[None, :, None], x[None, None, :]][axis]
    d.T = np.broadcast_to(1.0 + g * prof, (n, n, n)).ravel().copy()
    p = FluidParams(2.0, 0.01, 0.01, 0.0)
    dt = 1e-3
    f = _run1(d, p, dt)
    for k in range(3):
        c = _centre(f[k], d.shape, 8)
        if k == axis:
            assert np.allclose(c, -(dt / 3) * p.v_T * g, rtol=1e-9)
        else:
            # zero up to the roundoff of the merged Picard layer's T
            assert np.abs(f[k]).max() < 1e-14
    assert np.allclose(f[3], d.T, rtol=0, atol=1e-12)  # T linear: unchanged


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_advection_sign_per_axis(axis):
    """Uniform velocity c along one axis (all walls move with it) and T linear
    along that axis: implicit central advection gives T - (dt/3) c gamma."""
    n = 24
    d = _box(n)
    vel = [d.vx, d.vy, d.vz][axis]
    vel[:] = 1.5
    g = 0.4
    x = np.arange(n) * d.dx
    prof = [x[:, None, None], x[None, :, None], x[None, None, :]][axis]
    d.T = np.broadcast_to(1.0 + g * prof, (n, n, n)).ravel().copy()
    p = FluidParams(0.0, 0.01, 0.01, 0.0)
    dt = 1e-3
    f = _run1(d, p, dt, G=2)
    assert np.allclose(f[axis], 1.5, rtol=0, atol=1e-13)
    c = _centre(f[3], d.shape, 8)
    want = _centre(d.T, d.shape, 8) - (dt / 3) * 1.5 * g
    assert np.allclose(c, want, rtol=1e-11)


@pytest.mark.parametrize("case", ["u_y", "u_z", "v_x", "w_x", "v_z", "w_y", "cross_xy"])
def test_dissipation_terms(case):
    """Linear shear fields leave the velocity unchanged along the sweeps that see
    them; T gains (dt/3) t_phi Phi_d per sweep, Phi_d of TimeLayer3D.h:554-588."""
    n = 24
    d = _box(n)
    x = np.arange(n) * d.dx
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    a, b = 0.8, 0.5
    fields = {"u": np.zeros_like(X), "v": np.zeros_like(X), "w": np.zeros_like(X)}
    if case == "u_y":
        fields["u"] = a * Y; phi = {"y": a * a}
    elif case == "u_z":
        fields["u"] = a * Z; phi = {"z": a * a}
    elif case == "v_x":
        fields["v"] = a * X; phi = {"x": a * a}
    elif case == "w_x":
        fields["w"] = a * X; phi = {"x": a * a}
    elif case == "v_z":
        fields["v"] = a * Z; phi = {"z": a * a}
    elif case == "w_y":
        fields["w"] = a * Y; phi = {"y": a * a}
    else:  # u = a y, v = b x: Phi_X = v_x^2 + v_x u_y, Phi_Y = u_y^2 + u_y v_x
        fields["u"] = a * Y
        fields["v"] = b * X
        phi = {"x": b * b + b * a, "y": a * a + a * b}
    d.vx[:] = fields["u"].ravel()
    d.vy[:] = fields["v"].ravel()
    d.vz[:] = fields["w"].ravel()
    p = FluidParams(0.0, 0.01, 0.01, 0.3)
    dt = 1e-3
    f = _run1(d, p, dt)
    total = sum(phi.values())
    got = _centre(f[3], d.shape, 8) - 1.0
    assert np.allclose(got, (dt / 3) * p.t_phi * total, rtol=1e-8, atol=1e-15)


def test_zero_state_stays_zero():
    d = domains.lid_sphere(16, radius=3)
    d.vx[:] = 0
    d.T[:] = 0
    o = oracle_lib.Oracle3D(d, "f64")
    o.create_segments()
    assert o.run(3)[0] == 0
    for a in o.fields():
        assert np.all(a == 0)


# ------------------------------------------------------- composed step ----
@pytest.mark.parametrize("seed,G,L,free", [(0, 1, 1, False), (1, 2, 1, True), (2, 2, 2, False), (3, 3, 1, True)])
def test_step_vs_dense_brute_force(seed, G, L, free):
    from brute_adi import brute_step

    rng = np.random.default_rng(seed)
    d = domains.random_domain((7, 6, 8), rng, 0.15, free_faces=free)
    # a non-trivial state
    N = d.n
    inner = d.type == 0
    st = [np.where(inner, rng.normal(0, 0.3, N), a) for a in (d.vx, d.vy, d.vz)]
    Tst = np.where(inner, 1.0 + rng.normal(0, 0.1, N), d.T)
    p = FluidParams(1.0, 0.02, 0.03, 0.2)
    dt = 2e-3
    o = oracle_lib.Oracle3D(d, "f64", params=p)
    o.create_segments()
    o.set_fields(*st, Tst)
    o.update_boundaries()
    r, _ = o.step(dt, G, L, False)
    assert r == 0
    want = brute_step(d, p, dt, G, L, [a.copy() for a in st] + [Tst.copy()])
    got = o.fields()
    for k in range(4):
        assert np.allclose(got[k], want[k], rtol=1e-11, atol=1e-12), k


def test_gpu_merge_order_vs_dense_brute_force():
    from brute_adi import brute_step

    rng = np.random.default_rng(9)
    d = domains.random_domain((7, 8, 6), rng, 0.1)
    p = FluidParams(1.0, 0.02, 0.03, 0.2)
    st = [np.where(d.type == 0, rng.normal(0, 0.3, d.n), a) for a in (d.vx, d.vy, d.vz)]
    o = oracle_lib.Oracle3D(d, "f64", params=p, merge_mode=1)
    o.create_segments()
    o.set_fields(*st, d.T)
    o.update_boundaries()
    o.step(2e-3, 2, 1, False)
    want = brute_step(d, p, 2e-3, 2, 1, [a.copy() for a in st] + [d.T.copy()], gpu_order=True)
    got = o.fields()
    for k in range(4):
        assert np.allclose(got[k], want[k], rtol=1e-11, atol=1e-12), k


def test_divergence_error_closed_form():
    """Linear field u = alpha x (others 0): each dual cell's flux divergence is
    alpha * h^3 exactly (8-corner averages), so the mean is |alpha| h^3."""
    n = 10
    d = _box(n)
    x = np.arange(n) * d.dx
    u = np.broadcast_to(0.5 * x[:, None, None], (n, n, n)).ravel()
    o = oracle_lib.Oracle3D(d, "f64")
    o.set_fields(u, np.zeros(d.n), np.zeros(d.n), d.T)
    assert np.isclose(o.div_error(), 0.5 * d.dx**3, rtol=1e-12)


# ------------------------------------------------------------- output view ----
def test_get_layer_closed_form():
    """GetLayer (Solver3D.cpp:21-25 -> FilterToArrays): output node (i, j, k)
    takes grid node (i*nx/ox, j*ny/oy, k*nz/oz) of the layer that held the
    state before the last step; NODE_OUT nodes read 99999.  An index-valued
    state makes every pick visible."""
    d = domains.box(9, 7, 8)
    t = d.type.reshape(d.shape)
    t[4, 3, 0] = domains.NODE_OUT  # a face node, not next to NODE_IN
    o = oracle_lib.Oracle3D(d, "f64")
    o.create_segments()
    ids = np.arange(d.n, dtype=np.float64)
    o.set_fields(ids, 2 * ids, 3 * ids, 4 * ids)
    o.update_boundaries()
    o.step(1e-3, 1, 1, False)  # the index state moves to the `next` buffer
    # that layer is the index state with the node values written on BOUND/VALVE
    # nodes by UpdateBoundaries
    bv = (d.type == domains.NODE_BOUND) | (d.type == domains.NODE_VALVE)
    U = np.where(bv, d.vx, ids)
    V = np.where(bv, d.vy, 2 * ids)
    T4 = np.where(bv, d.T, 4 * ids)
    for ox, oy, oz in ((9, 7, 8), (4, 3, 5), (12, 10, 9)):
        vel, T = o.get_layer(ox, oy, oz)
        i = np.arange(ox)[:, None, None] * d.nx // ox
        j = np.arange(oy)[None, :, None] * d.ny // oy
        k = np.arange(oz)[None, None, :] * d.nz // oz
        src = ((i * d.ny + j) * d.nz + k).ravel()
        out = d.type[src] == domains.NODE_OUT
        assert out.any() or (ox, oy, oz) != (9, 7, 8)
        assert np.array_equal(vel[:, 0], np.where(out, 99999.0, U[src]).astype(np.float32))
        assert np.array_equal(vel[:, 1], np.where(out, 99999.0, V[src]).astype(np.float32))
        assert np.array_equal(T, np.where(out, 99999.0, T4[src]))
