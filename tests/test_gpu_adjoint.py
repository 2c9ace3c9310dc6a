"""GPU reverse mode: the HIP adjoint against the oracle's adjoint and against
central finite differences of the HIP forward.

Bars: FP64 gradients within 1e-9 relative L2 of the FP64 oracle adjoint (node
values, initial state, each FluidParams component); FP32 gradients within
1e-3 of it; FP64 directional derivatives within 1e-5 of central differences
(step 1e-3) on a channel with a cylinder; the torch.autograd binding
returns the same gradients as the C ABI.
"""
import numpy as np
import pytest

import adjoint_cases as ac
import oracle_lib
from cmcx import domains
from cmcx.solver import make_batch, make_solver
from gpu_util import rel_l2, say

pytestmark = pytest.mark.gpu

PLAIN, FUSED = 0, 1
STEPS, EVERY = 5, 2  # checkpoint intervals of 2, 2 and 1 steps


def _oracle_grads(d, seed):
    o = oracle_lib.Oracle3D(d, "f64")
    o.create_segments()
    o.adj_record(STEPS, ac.DT, ac.G)
    gb, cb, pb = o.adj_backward(seed)
    o.close()
    return gb, cb, pb


@pytest.fixture(scope="module")
def small():
    d = ac.small_domain()
    rng = np.random.default_rng(23)
    seed = rng.normal(size=(4, d.n)) * (d.type == domains.NODE_IN)
    return d, seed, _oracle_grads(d, seed)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("path", [PLAIN, FUSED])
def test_adjoint_matches_oracle(small, dtype, path):
    d, seed, (ogb, ocb, opb) = small
    say(f"[adjoint vs oracle] {dtype} path={path}")
    s = make_solver(d, dtype, path=path)
    before = s.fields()
    s.adjoint_forward(STEPS, ac.DT, ac.G, EVERY)
    after = s.fields()
    gb, sb, pb = s.adjoint_backward(seed)
    # the backward pass leaves the final state in place, and a second backward
    # pass (which recomputes the last interval the forward had recorded) agrees
    for x, y in zip(after, s.fields()):
        assert np.array_equal(x, y)
    gb2, sb2, pb2 = s.adjoint_backward(seed)
    assert rel_l2(gb2, gb) < 1e-12 and rel_l2(sb2, sb) < 1e-12
    assert not np.array_equal(before[0], after[0])
    tol = 1e-9 if dtype == "f64" else 1e-3
    eg, es = rel_l2(gb, ogb), rel_l2(sb, ocb)
    say(f"  node values {eg:.2e}  initial state {es:.2e}")
    assert eg < tol and es < tol
    for q in range(4):
        e = abs(pb[q] - opb[q]) / max(abs(opb[q]), 1e-30)
        say(f"  param {q}: {pb[q]:.10e} vs {opb[q]:.10e} ({e:.2e})")
        assert e < tol, q
    s.close()


def test_adjoint_long_lines():
    """Lines longer than 128 nodes along X and Y (the transposed solve then uses
    chunks of 8 rows), FP64 against the oracle's adjoint."""
    rng = np.random.default_rng(31)
    for shape in ((140, 9, 10), (9, 140, 10)):
        d = domains.random_domain(shape, rng, 0.08, free_faces=True)
        d.re, d.dt = 50.0, ac.DT
        seed = rng.normal(size=(4, d.n)) * (d.type == domains.NODE_IN)
        say(f"[adjoint long lines] {shape}")
        ogb, ocb, opb = _oracle_grads(d, seed)
        s = make_solver(d, "f64", path=FUSED)
        s.adjoint_forward(STEPS, ac.DT, ac.G, EVERY)
        gb, sb, pb = s.adjoint_backward(seed)
        eg, es = rel_l2(gb, ogb), rel_l2(sb, ocb)
        say(f"  node values {eg:.2e}  initial state {es:.2e}")
        assert eg < 1e-9 and es < 1e-9
        assert np.allclose(pb, opb, rtol=1e-9, atol=0)
        s.close()


def _channel(n=36, steps=12):
    # dt 1e-2: inlet changes reach the probe next to the inlet within the run
    d, control = domains.cylinder_channel(n=n, radius=5.0, seed=1, steps=steps, dt=1e-2, num_global=4)
    return d, control


def test_adjoint_matches_finite_differences():
    d, control = _channel()
    mask = ac.probe_mask(d.shape, size=8, x0=2)
    say(f"[adjoint vs FD] channel {d.shape}, {d.steps} steps")
    s = make_solver(d, "f64", path=FUSED)
    init = s.fields()
    state0 = [init[0], init[1], init[2], init[4]]
    ids = np.arange(d.ny * d.nz)  # inlet face x = 0

    def set_inlet(c):
        vx = d.vx.copy()
        vx[ids] = np.kron(c, np.ones((3, 3)))[:d.ny, :d.nz].ravel()
        return vx

    def loss(c, nu):
        p = d.params()
        p.v_vis = nu
        s.set_params(p)
        s.set_boundary(d.bc_vel, d.bc_temp, set_inlet(c), d.vy, d.vz, d.T)
        s.set_fields(*state0)
        for _ in range(d.steps):  # the recorded rollout evaluates no divergence either
            s.UpdateBoundaries()
            s.TimeStep(d.dt, d.num_global, 1, False)
        return ac.probe_ke(s.fields(), mask)[0]

    nu0 = d.params().v_vis
    # adjoint at (control, nu0)
    s.set_boundary(d.bc_vel, d.bc_temp, set_inlet(control), d.vy, d.vz, d.T)
    s.set_fields(*state0)
    s.adjoint_forward(d.steps, d.dt, d.num_global, 4)
    L0, seed = ac.probe_ke(s.fields(), mask)
    say(f"  loss {L0:.6e}")
    assert L0 > 1e-3
    gb, _, pb = s.adjoint_backward(seed)
    gc = gb[0][ids].reshape(d.ny, d.nz)
    gctl = np.zeros_like(control)
    for J in range(control.shape[0]):
        for K in range(control.shape[1]):
            gctl[J, K] = gc[3 * J:3 * J + 3, 3 * K:3 * K + 3].sum()
    rng = np.random.default_rng(4)
    for trial in range(3):
        dc = rng.normal(size=control.shape)
        dn = rng.normal() * nu0
        eps = 1e-3
        fd = (loss(control + eps * dc, nu0 + eps * dn) - loss(control - eps * dc, nu0 - eps * dn)) / (2 * eps)
        ad = float(np.sum(gctl * dc) + pb[1] * dn)
        say(f"  direction {trial}: FD {fd:.10e} adjoint {ad:.10e} ({abs(fd - ad) / abs(fd):.2e})")
        assert abs(fd - ad) <= 1e-5 * abs(fd)
    s.close()


def test_torch_binding_matches_abi():
    torch = pytest.importorskip("torch")
    from cmcx.autograd import channel_rollout

    d, control = _channel(n=24, steps=6)
    mask = ac.probe_mask(d.shape, size=6, x0=2)
    say("[torch binding]")
    s = make_solver(d, "f64", path=FUSED)
    c = torch.tensor(control, requires_grad=True)
    nu = torch.tensor(d.params().v_vis, dtype=torch.float64, requires_grad=True)
    final = channel_rollout(s, d, c, nu, d.steps, every=3)
    m = torch.from_numpy(mask)
    loss = 0.5 * (final[0][m] ** 2 + final[1][m] ** 2 + final[2][m] ** 2).mean()
    loss.backward()
    # the same through the C ABI
    L0, seed = ac.probe_ke([final[k].detach().numpy() for k in range(3)], mask)
    assert abs(L0 - float(loss.detach())) <= 1e-12 * abs(L0)
    gb, _, pb = s.adjoint_backward(seed)
    ids = np.arange(d.ny * d.nz)
    gc = gb[0][ids].reshape(d.ny, d.nz)
    want = np.array([[gc[3 * J:3 * J + 3, 3 * K:3 * K + 3].sum() for K in range(control.shape[1])]
                     for J in range(control.shape[0])])
    assert rel_l2(c.grad.numpy(), want) < 1e-12
    assert abs(float(nu.grad) - pb[1]) <= 1e-12 * abs(pb[1])
    s.close()


def test_adjoint_rejects_batch():
    doms = domains.batch_boxes(count=2, seed=3, edges=(12, 16))
    s = make_batch(doms, "f32")
    with pytest.raises(Exception):
        s.adjoint_forward(2, 1e-3, 2, 1)
    s.close()


def test_recording_belongs_to_its_rollout():
    """A solver holds one recording: a second rollout recorded before the first
    one's backward, or new parameters or node values, make the first backward
    raise instead of returning another rollout's gradient."""
    torch = pytest.importorskip("torch")
    from cmcx._lib import CmcxError
    from cmcx.autograd import channel_rollout

    d, control = _channel(n=24, steps=4)
    mask = torch.from_numpy(ac.probe_mask(d.shape, size=6, x0=2))
    s = make_solver(d, "f64", path=FUSED)
    say("[recording generation]")
    c = torch.tensor(control, requires_grad=True)
    nu = torch.tensor(d.params().v_vis, dtype=torch.float64, requires_grad=True)
    f1 = channel_rollout(s, d, c, nu, d.steps, every=2)
    g1 = s.adjoint_generation()
    assert g1 != 0
    f2 = channel_rollout(s, d, c * 1.01, nu, d.steps, every=2)
    assert s.adjoint_generation() not in (0, g1)
    with pytest.raises(RuntimeError, match="no longer this rollout"):
        (f1[0][mask] ** 2).sum().backward()
    (f2[0][mask] ** 2).sum().backward()  # the live recording still works
    assert c.grad is not None and torch.isfinite(c.grad).all()
    # parameter and node-value changes clear the recording at the C ABI too
    s.adjoint_forward(2, d.dt, d.num_global, 1)
    s.set_params(d.params())
    assert s.adjoint_generation() == 0
    with pytest.raises(CmcxError, match="set_params"):
        s.adjoint_backward(None)
    s.adjoint_forward(2, d.dt, d.num_global, 1)
    s.set_boundary(d.bc_vel, d.bc_temp, d.vx, d.vy, d.vz, d.T)
    with pytest.raises(CmcxError, match="set_boundary"):
        s.adjoint_backward(None)
    s.close()


def test_backward_device_buffers_match_host():
    """cmcx_adjoint_backward_device (seed and gradients in device memory) gives
    the host call's gradients bit for bit, over a rollout of several checkpoint
    intervals (recompute and reverse overlapped on two streams)."""
    import ctypes

    from cmcx._lib import lib

    lib()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    d, control = _channel(n=24, steps=7)
    mask = ac.probe_mask(d.shape, size=6, x0=2)
    say("[adjoint device buffers]")
    for dtype in ("f32", "f64"):
        s = make_solver(d, dtype, path=FUSED)
        s.adjoint_forward(d.steps, d.dt, d.num_global, 2)  # intervals 2, 2, 2, 1
        _, seed = ac.probe_ke(s.fields(), mask)
        gh, sh, ph = s.adjoint_backward(seed)
        nbytes = d.n * np.dtype(s.np_dtype).itemsize
        bufs = []
        for _ in range(12):
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), nbytes) == 0
            bufs.append(p.value)
        for f in range(4):
            a = np.ascontiguousarray(seed[f], dtype=s.np_dtype)
            assert hip.hipMemcpy(bufs[f], a.ctypes.data, nbytes, 1) == 0
        pd = s.adjoint_backward_device(bufs[0:4], bufs[4:8], bufs[8:12])
        for f in range(4):
            g = np.empty(d.n, s.np_dtype)
            assert hip.hipMemcpy(g.ctypes.data, bufs[4 + f], nbytes, 2) == 0
            assert np.array_equal(g, gh[f]), (dtype, f)
            assert hip.hipMemcpy(g.ctypes.data, bufs[8 + f], nbytes, 2) == 0
            assert np.array_equal(g, sh[f]), (dtype, f)
        assert np.array_equal(pd, ph)
        for p in bufs:
            hip.hipFree(p)
        s.close()
