"""GPU parity: the HIP path through the C ABI against the oracle.

Bars: segment lists and endpoint records bit-exact; FP64 fields within 1e-11
relative L2 of the FP64 oracle; FP32 fields within 1e-5 relative L2 of the
FP32 oracle (the reference's own precision) and of the FP64 oracle.
"""
import numpy as np
import pytest

import oracle_lib
from cmcx import domains
from cmcx.solver import make_solver
from gpu_util import rel_l2, rel_l2_scaled, say, vel_norm

pytestmark = pytest.mark.gpu

PLAIN, FUSED = 0, 1


def _free_wall():
    """Box with a one-node interior wall with free BCs: two X segments share it."""
    d = domains.random_domain((20, 16, 18), np.random.default_rng(11), 0.0)
    t = d.type.reshape(d.shape)
    m = np.zeros(d.shape, bool)
    m[10, 3:13, 3:15] = True
    t[m] = domains.NODE_BOUND
    d.bc_vel.reshape(d.shape)[m] = domains.BC_FREE
    d.bc_temp.reshape(d.shape)[m] = domains.BC_FREE
    for a in (d.vx, d.vy, d.vz):
        a.reshape(d.shape)[m] = 0.0
    return d


def _cases():
    rng = np.random.default_rng(7)
    return {
        "free_wall_20x16x18": _free_wall(),
        "lid_sphere_24": domains.lid_sphere(24, radius=5.0, dt=1e-3),
        "random_18x20x22": domains.random_domain((18, 20, 22), rng, 0.12),
        "random_free_17x13x21": domains.random_domain((17, 13, 21), rng, 0.10, free_faces=True),
        "ellipsoids_40": domains.ellipsoids(40, count=6, axis_lo=3, axis_hi=8, seed=3, dt=5e-4),
    }


CASES = _cases()


@pytest.mark.parametrize("name", list(CASES))
def test_segments_bit_exact(name):
    d = CASES[name]
    say(f"[segments] {name} {d.shape}")
    ref = oracle_lib.Oracle3D(d, "f64")
    ref.create_segments()
    s = make_solver(d, "f64")
    N = d.shape
    for a in range(3):
        got, want = s.segments(a), ref.get_segments(a)
        assert got.shape == want.shape, (a, got.shape, want.shape)
        assert np.array_equal(got, want)
        flags, vals = s.segment_nodes(a)
        if len(want):
            first = (want[:, 0] * N[1] + want[:, 1]) * N[2] + want[:, 2]
            last = (want[:, 3] * N[1] + want[:, 4]) * N[2] + want[:, 5]
            for e, ids in enumerate((first, last)):
                assert np.array_equal(flags[:, 3 * e], d.type[ids])
                assert np.array_equal(flags[:, 3 * e + 1], d.bc_vel[ids])
                assert np.array_equal(flags[:, 3 * e + 2], d.bc_temp[ids])
                assert np.array_equal(vals[:, 4 * e + 0], d.vx[ids])
                assert np.array_equal(vals[:, 4 * e + 3], d.T[ids])
    s.close()


@pytest.mark.parametrize("nb", [2, 3])
def test_segments_blocked_order(nb):
    d = CASES["random_18x20x22"]
    say(f"[segments nblockz={nb}]")
    s = make_solver(d, "f64", nblockz=nb)
    ref = oracle_lib.Oracle3D(d, "f64", nblockz=nb)
    ref.create_segments()
    for a in range(3):
        assert np.array_equal(s.segments(a), ref.get_segments(a))
    s.close()


def _run_pair(d, dtype, path, steps, G=2, L=1, merge_order=0, oracle_dtype=None, variant=None):
    ref = oracle_lib.Oracle3D(d, oracle_dtype or dtype, merge_mode=merge_order)
    ref.create_segments()
    s = make_solver(d, dtype, path=path, merge_order=merge_order, variant=variant)
    errs = []
    for i in range(steps):
        ref.update_boundaries()
        s.UpdateBoundaries()
        ce = (i % 10 == 0) or i == steps - 1
        r, e_ref = ref.step(d.dt, G, L, ce)
        assert r == 0
        e = s.TimeStep(d.dt, G, L, ce)
        errs.append((e, e_ref))
    return s, ref, errs


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("path", [PLAIN, FUSED])
def test_step_fp64(name, path):
    d = CASES[name]
    say(f"[fp64] {name} path={path}")
    s, ref, errs = _run_pair(d, "f64", path, steps=3)
    u, v, w, p, T = s.fields()
    ru, rv, rw, rT = ref.fields()
    vs = vel_norm(ru, rv, rw)
    for nm, a, b, sc in (("u", u, ru, vs), ("v", v, rv, vs), ("w", w, rw, vs), ("T", T, rT, 0.0),
                         ("p", p, rT * d.params().v_T, 0.0)):
        e = rel_l2_scaled(a, b, sc)
        say(f"  {nm}: rel L2 vs FP64 oracle {e:.3e}")
        assert e < 1e-11, nm
    for e, e_ref in errs:
        assert abs(e - e_ref) <= 1e-12 * max(abs(e_ref), 1e-30) + 1e-300
    s.close()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("path", [PLAIN, FUSED])
def test_step_fp32(name, path):
    d = CASES[name]
    say(f"[fp32] {name} path={path}")
    s, ref, _ = _run_pair(d, "f32", path, steps=3)
    ref64 = oracle_lib.Oracle3D(d, "f64")
    ref64.create_segments()
    ref64.run(3, G=2)
    u, v, w, p, T = s.fields()
    f32 = ref.fields()
    f64 = ref64.fields()
    vs = vel_norm(*f64[:3])
    for k, nm in enumerate("uvwT"):
        a = (u, v, w, T)[k]
        sc = vs if k < 3 else 0.0
        e32 = rel_l2_scaled(a, f32[k], sc)
        e64 = rel_l2_scaled(a, f64[k], sc)
        say(f"  {nm}: rel L2 vs FP32 oracle {e32:.3e}, vs FP64 oracle {e64:.3e}")
        assert e32 < 1e-5 and e64 < 1e-5, nm
    s.close()


def test_gpu_merge_order():
    d = CASES["lid_sphere_24"]
    say("[merge order: reference CUDA]")
    s, ref, _ = _run_pair(d, "f64", PLAIN, steps=2, merge_order=1)
    got = s.fields()
    want = ref.fields()
    vs = vel_norm(*want[:3])
    for k in range(3):
        assert rel_l2_scaled(got[k], want[k], vs) < 1e-11
    assert rel_l2(got[4], want[3]) < 1e-11
    s.close()


@pytest.mark.parametrize("path", [PLAIN, FUSED])
def test_local_iterations(path):
    d = CASES["random_18x20x22"]
    say(f"[num_local=2] path={path}")
    s, ref, _ = _run_pair(d, "f64", path, steps=2, G=2, L=2)
    got = s.fields()
    want = ref.fields()
    vs = vel_norm(*want[:3])
    for k in range(3):
        assert rel_l2_scaled(got[k], want[k], vs) < 1e-11
    s.close()


@pytest.mark.parametrize("path", [PLAIN, FUSED])
def test_get_layer_matches_reference_view(path):
    d = CASES["lid_sphere_24"]
    say(f"[get_layer] path={path}")
    s, ref, _ = _run_pair(d, "f64", path, steps=2)
    v, T = s.GetLayer(12, 12, 12)
    rv, rT = ref.get_layer(12, 12, 12)
    assert np.allclose(v, rv, rtol=1e-6, atol=1e-7)
    assert np.allclose(T, rT, rtol=1e-12, atol=0)
    s.close()


def test_divergence_guard_keeps_state():
    from cmcx._lib import DivergenceError

    d = CASES["lid_sphere_24"]
    say("[divergence guard]")
    s = make_solver(d, "f64")
    rng = np.random.default_rng(1)
    u0 = np.where(d.type == 0, 100.0 * rng.standard_normal(d.n), d.vx)
    s.set_fields(u0, d.vy, d.vz, d.T)
    s.UpdateBoundaries()
    before = s.fields()
    with pytest.raises(DivergenceError):
        s.TimeStep(d.dt, 2, 1, True)  # mean |div| of a noise field far above 0.01
    after = s.fields()
    for a, b in zip(before, after):
        assert np.array_equal(a, b)  # no swap after the guard, as in the reference
    s.close()


def test_mask_rejected():
    from cmcx._lib import CmcxError

    d = domains.box(8, 8, 8)
    d.type[0] = 0  # NODE_IN on a corner
    say("[mask check]")
    with pytest.raises(CmcxError):
        make_solver(d, "f32")


@pytest.mark.parametrize("variant", list(range(10)))
@pytest.mark.parametrize("n", [32, 48, 64, 96, 128])
def test_fused_line_lengths(n, variant):
    """Fused kernels at every line length of the benchmark workloads (chunk and
    tile padding), every tile variant (the chunk-4 ones fit all these lengths),
    FP64 against the FP64 oracle, FP32 against the FP64 oracle."""
    d = domains.lid_sphere(n, radius=n / 6.0, dt=1e-3)
    say(f"[fused length] {n}^3 variant {variant}")
    for dtype, tol in (("f64", 1e-11), ("f32", 1e-5)):
        s, ref, _ = _run_pair(d, dtype, FUSED, steps=2, oracle_dtype="f64", variant=variant)
        got = s.fields()
        want = ref.fields()
        vs = vel_norm(*want[:3])
        for k in range(3):
            e = rel_l2_scaled(got[k], want[k], vs)
            say(f"  {dtype} field {k}: {e:.3e}")
            assert e < tol
        assert rel_l2(got[4], want[3]) < tol
        s.close()


def test_fused_anisotropic_grid():
    d = domains.random_domain((40, 24, 70), np.random.default_rng(5), 0.1, free_faces=True)
    say("[fused anisotropic] 40x24x70")
    s, ref, _ = _run_pair(d, "f64", FUSED, steps=2)
    got = s.fields()
    want = ref.fields()
    vs = vel_norm(*want[:3])
    for k in range(3):
        assert rel_l2_scaled(got[k], want[k], vs) < 1e-11
    assert rel_l2(got[4], want[3]) < 1e-11
    s.close()


@pytest.mark.parametrize("variant", list(range(10)))
def test_fused_variant_ellipsoids(variant):
    d = CASES["ellipsoids_40"]
    say(f"[fused variant] ellipsoids_40 variant {variant}")
    s, ref, _ = _run_pair(d, "f64", FUSED, steps=3, variant=variant)
    got = s.fields()
    want = ref.fields()
    vs = vel_norm(*want[:3])
    for k in range(3):
        assert rel_l2_scaled(got[k], want[k], vs) < 1e-11
    assert rel_l2(got[4], want[3]) < 1e-11
    s.close()


def test_autotune_keeps_state_and_parity():
    d = CASES["random_free_17x13x21"]
    say("[autotune]")
    ref = oracle_lib.Oracle3D(d, "f64")
    ref.create_segments()
    s = make_solver(d, "f64", path=FUSED)
    before = s.fields()
    choice = s.autotune(2)
    assert len(choice) == 3 and all(0 <= c < 10 for c in choice[:2]) and 0 <= choice[2] < 8
    after = s.fields()
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    for i in range(2):
        ref.update_boundaries()
        ref.step(d.dt, 2, 1, True)
        s.UpdateBoundaries()
        s.TimeStep(d.dt, 2, 1, True)
    got = s.fields()
    want = ref.fields()
    vs = vel_norm(*want[:3])
    for k in range(3):
        assert rel_l2_scaled(got[k], want[k], vs) < 1e-11
    s.close()
