"""Pins the oracle's discrete adjoint of the time step (CPU, no GPU).

Central finite differences of the FP64 oracle's forward map, along random
directions in every input group the adjoint differentiates:
  * the node values of BOUND/VALVE nodes (inlet control, lid, walls),
  * the initial state on NODE_IN nodes,
  * each of the four FluidParams (v_T, v_vis, t_vis, t_phi),
against the adjoint's inner product <grad, direction>.  The loss is a random
linear functional of the final state plus its kinetic energy, so that the
seed is state dependent.  A planted sign error in any scatter of the adjoint
(pressure, dissipation, advecting velocity, merge weights, end rows) moves the
inner product far outside the tolerance; the last test plants one in the
oracle's output and checks that the comparison sees it.
"""
import numpy as np
import pytest

import oracle_lib
from cmcx import domains
from cmcx.params import FluidParams

STEPS, G, DT = 3, 2, 4e-3


def _domain():
    rng = np.random.default_rng(5)
    d = domains.random_domain((9, 8, 7), rng, 0.08, free_faces=True)
    sh = d.shape
    t = d.type.reshape(sh)
    # inlet x = 0: VALVE, Dirichlet, non-uniform u
    t[0] = domains.NODE_VALVE
    d.vx.reshape(sh)[0] = 0.5 + rng.random(sh[1:])
    d.vy.reshape(sh)[0] = 0.1 * rng.random(sh[1:])
    d.T.reshape(sh)[0] = 1.0 + 0.2 * rng.random(sh[1:])
    d.re = 50.0
    d.dt = DT
    return d


def _params(d, dp=None, eps=0.0):
    p = d.params()
    v = np.array([p.v_T, p.v_vis, p.t_vis, p.t_phi], np.float64)
    if dp is not None:
        v = v + eps * dp
    return FluidParams(*v)


def _forward(d, params=None, state=None):
    o = oracle_lib.Oracle3D(d, "f64", params=params)
    o.create_segments()
    if state is not None:
        o.set_fields(*state)
    for _ in range(STEPS):
        o.update_boundaries()
        r, _ = o.step(DT, G, 1, False)
        assert r == 0
    f = np.array(o.fields())
    o.close()
    return f


def _loss(f, W):
    return float(np.sum(W * f) + 0.5 * np.sum(f[:3] ** 2 * (W[0] > 0)))


def _seed(f, W):
    s = W.copy()
    s[:3] += f[:3] * (W[0] > 0)
    return s


@pytest.fixture(scope="module")
def setup():
    d = _domain()
    rng = np.random.default_rng(17)
    inn = d.type == domains.NODE_IN
    W = rng.normal(size=(4, d.n)) * inn
    o = oracle_lib.Oracle3D(d, "f64")
    o.create_segments()
    state0 = np.array(o.fields())
    o.adj_record(STEPS, DT, G)
    f = np.array(o.fields())
    gb, cb, pb = o.adj_backward(_seed(f, W))
    o.close()
    return d, W, inn, state0, gb, cb, pb, rng


def _fd(fun, eps):
    return (fun(eps) - fun(-eps)) / (2 * eps)


def test_adjoint_node_values(setup):
    d, W, inn, state0, gb, cb, pb, rng = setup
    bv = (d.type == domains.NODE_BOUND) | (d.type == domains.NODE_VALVE)
    dg = rng.normal(size=(4, d.n)) * bv

    def L(eps):
        e = d.copy()
        e.vx, e.vy, e.vz, e.T = (d.vx + eps * dg[0], d.vy + eps * dg[1], d.vz + eps * dg[2],
                                 d.T + eps * dg[3])
        return _loss(_forward(e), W)

    fd = _fd(L, 1e-6)
    ad = float(np.sum(gb * dg))
    assert abs(fd - ad) <= 1e-7 * max(abs(fd), 1e-12), (fd, ad)


def test_adjoint_initial_state(setup):
    d, W, inn, state0, gb, cb, pb, rng = setup
    ds = rng.normal(size=(4, d.n)) * inn

    def L(eps):
        return _loss(_forward(d, state=state0 + eps * ds), W)

    fd = _fd(L, 1e-6)
    ad = float(np.sum(cb * ds))
    assert abs(fd - ad) <= 1e-7 * max(abs(fd), 1e-12), (fd, ad)


@pytest.mark.parametrize("q", [0, 1, 2, 3])
def test_adjoint_params(setup, q):
    d, W, inn, state0, gb, cb, pb, rng = setup
    dp = np.zeros(4)
    dp[q] = 1.0
    base = _params(d)
    scale = [base.v_T, base.v_vis, base.t_vis, base.t_phi][q]

    def L(eps):
        return _loss(_forward(d, params=_params(d, dp, eps)), W)

    # a relative step of 1e-3: the truncation error (1e-9 here) stays below the
    # roundoff of smaller steps
    fd = _fd(L, 1e-3 * scale)
    assert abs(fd - pb[q]) <= 1e-7 * max(abs(fd), 1e-12), (q, fd, pb[q])


def test_planted_error_is_seen(setup):
    # the comparison of test_adjoint_node_values with a sign error planted in
    # the adjoint's inlet gradient must fail
    d, W, inn, state0, gb, cb, pb, rng = setup
    bad = gb.copy()
    inlet = d.type.reshape(d.shape).copy()
    inlet[1:] = 255
    bad[0, inlet.ravel() == domains.NODE_VALVE] *= -1
    bv = (d.type == domains.NODE_BOUND) | (d.type == domains.NODE_VALVE)
    dg = rng.normal(size=(4, d.n)) * bv

    def L(eps):
        e = d.copy()
        e.vx = d.vx + eps * dg[0]
        e.vy, e.vz, e.T = d.vy + eps * dg[1], d.vz + eps * dg[2], d.T + eps * dg[3]
        return _loss(_forward(e), W)

    fd = _fd(L, 1e-6)
    assert abs(fd - float(np.sum(bad * dg))) > 1e-3 * abs(fd)
