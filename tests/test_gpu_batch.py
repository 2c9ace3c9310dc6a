"""GPU: several grids advanced together, and steps replayed as HIP graphs.

Batch: every grid of a batch is checked against its own oracle run (FP64
within 1e-11 relative L2 of the FP64 oracle, FP32 within 1e-5 of the FP32 and
FP64 oracles); a batch of one is bit-identical to the single-grid solver.
Graphs: the replayed step is bit-identical to the directly launched one after
every step, with host reads (eager work and synchronisation) between replays.
"""
import numpy as np
import pytest

import oracle_lib
from cmcx import domains
from cmcx.solver import make_batch, make_solver
from gpu_util import rel_l2_scaled, say, vel_norm

pytestmark = pytest.mark.gpu

PLAIN, FUSED = 0, 1


def _boxes(count=5, seed=1):
    return domains.batch_boxes(count=count, seed=seed, edges=(12, 16, 20, 28), dt=1e-3)


def _free_wall():
    d = domains.random_domain((20, 16, 18), np.random.default_rng(11), 0.0)
    t = d.type.reshape(d.shape)
    m = np.zeros(d.shape, bool)
    m[10, 3:13, 3:15] = True
    t[m] = domains.NODE_BOUND
    d.bc_vel.reshape(d.shape)[m] = domains.BC_FREE
    d.bc_temp.reshape(d.shape)[m] = domains.BC_FREE
    for a in (d.vx, d.vy, d.vz):
        a.reshape(d.shape)[m] = 0.0
    d.dt = 1e-3
    return d


def _oracle_fields(d, dtype, steps, G):
    o = oracle_lib.Oracle3D(d, dtype)
    o.create_segments()
    errs = []
    for i in range(steps):
        o.update_boundaries()
        r, e = o.step(d.dt, G, 1, True)
        assert r == 0
        errs.append(e)
    f = o.fields()
    o.close()
    return f, errs


def _check_batch(doms, dtype, path, steps=3, G=2, variant=None):
    s = make_batch(doms, dtype, path=path)
    assert len(s) == len(doms)
    if variant is not None:
        for a in (0, 1):
            s.set_tile_variant(a, variant)
    dt = doms[0].dt
    last = None
    for i in range(steps):
        s.UpdateBoundaries()
        s.TimeStep(dt, G, 1, True)
        last = s.grid_errors()
    tol = 1e-11 if dtype == "f64" else 1e-5
    for k, d in enumerate(doms):
        u, v, w, p, T = s.fields_of(k)
        refs = [_oracle_fields(d, "f64", steps, G)]
        if dtype == "f32":
            refs.append(_oracle_fields(d, "f32", steps, G))
        for (rf, rerr) in refs:
            vs = vel_norm(*rf[:3])
            for j, nm in enumerate("uvwT"):
                e = rel_l2_scaled((u, v, w, T)[j], rf[j], vs if j < 3 else 0.0)
                assert e < tol, (k, d.name, nm, e)
            if dtype == "f64":
                assert abs(last[k] - rerr[-1]) <= 1e-9 * max(abs(rerr[-1]), 1e-30), (k, last[k], rerr[-1])
        say(f"  grid {k} {d.shape}: ok")
    s.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("path", [PLAIN, FUSED])
def test_batch_matches_oracle(dtype, path):
    doms = _boxes()
    say(f"[batch] {len(doms)} grids {[d.shape for d in doms]} {dtype} path={path}")
    _check_batch(doms, dtype, path)


@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("variant", [6, 7, 9])
def test_batch_lds_parked_variants(variant, dtype):
    """X and Y sweeps with the Picard layer parked in LDS on grids of mixed
    sizes in one launch (the work-group is sized for the longest line; the
    shorter lines end in identity chunks)."""
    doms = _boxes()
    say(f"[batch variant {variant}] {dtype}")
    _check_batch(doms, dtype, FUSED, variant=variant)


def test_batch_with_shared_free_node():
    # one grid needs the per-segment kernel on X: that axis runs it for every grid
    doms = [_free_wall()] + _boxes(3, seed=4)
    say("[batch with a shared free node]")
    _check_batch(doms, "f64", FUSED)


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_batch_of_one_is_single(dtype):
    d = domains.ellipsoids(40, count=6, axis_lo=3, axis_hi=8, seed=3, dt=5e-4)
    say(f"[batch of one] {dtype}")
    a = make_batch([d], dtype, path=FUSED)
    b = make_solver(d, dtype, path=FUSED)
    for i in range(3):
        for s in (a, b):
            s.UpdateBoundaries()
            s.TimeStep(d.dt, 2, 1, False)
    for x, y in zip(a.fields_of(0), b.fields()):
        assert np.array_equal(x, y)
    a.close()
    b.close()


@pytest.mark.parametrize("path", [PLAIN, FUSED])
def test_graph_replay_bit_identical(path):
    d = domains.ellipsoids(40, count=6, axis_lo=3, axis_hi=8, seed=3, dt=5e-4)
    say(f"[graph] path={path}")
    g = make_solver(d, "f32", path=path)
    e = make_solver(d, "f32", path=path)
    g.set_graph(True)
    for i in range(6):  # each buffer parity's graph is replayed three times
        for s in (g, e):
            s.UpdateBoundaries()
            err = s.TimeStep(d.dt, 2, 1, i % 2 == 0)
        for x, y in zip(g.fields(), e.fields()):
            assert np.array_equal(x, y), i
    # whole-loop graphs (boundary fill included), then back to direct launches
    g.run(5, d.dt, 2, 1)
    e.run(5, d.dt, 2, 1)
    for x, y in zip(g.fields(), e.fields()):
        assert np.array_equal(x, y)
    g.set_graph(False)
    for s in (g, e):
        s.UpdateBoundaries()
        s.TimeStep(d.dt, 2, 1, False)
    for x, y in zip(g.fields(), e.fields()):
        assert np.array_equal(x, y)
    g.close()
    e.close()


def test_graph_batch():
    doms = _boxes(4, seed=9)
    say("[graph batch]")
    g = make_batch(doms, "f32")
    e = make_batch(doms, "f32")
    g.set_graph(True)
    for i in range(4):
        for s in (g, e):
            s.run(2, doms[0].dt, 2, 1)
        for k in range(len(doms)):
            for x, y in zip(g.fields_of(k), e.fields_of(k)):
                assert np.array_equal(x, y), (i, k)
    g.close()
    e.close()
