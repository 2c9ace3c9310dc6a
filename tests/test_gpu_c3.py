"""The throughput workload at full size: 256^3 with 40 seeded ellipsoids removed
(seed 0), inlet u = 1 at x = 0, BC_FREE outlet at x = 255, Re 1000, dt 2e-4,
4 Picard iterations -- the grid, kernels and tile variants the headline
cell-updates/s figure is measured on.

* Segment lists of all three axes bit-exact against the oracle.
* 5 steps on the GPU against 5 steps of the FP64 oracle (and of the FP32
  oracle) from the same initial state: FP32 with tile variants 6 on every axis
  (128-byte X/Y tiles of 32 full z lines with the Picard layer parked in the
  whole 160 KB of LDS, and the Z sweep with 64 lanes of 4 rows per z line and
  its Picard layer parked in LDS) and with the default variants (128-byte X/Y
  tiles reading the Picard layer again for the merge), the plain per-segment
  FP32 path, and FP64.  Errors are reported per component relative to that
  component's own signal (u, v, w; T - baseT; p - v_T baseT) beside the
  velocity-norm-scaled figure.  Bars: FP64 within 1e-11 own-relative L2; FP32
  within 1e-5 velocity-scaled relative L2 of the FP64 oracle, and each
  component's own-relative error at most twice the FP32 oracle's own error
  against the FP64 oracle; the plain FP32 path bit-identical to the FP32
  oracle (same operation order, no contraction).
* The configured 50 steps (boundary fill, error every 10th step) with property
  checks: finite fields, every no-slip BOUND/VALVE node equal to its node
  value, a divergence error under the guard.
"""
import numpy as np
import pytest

import oracle_lib
from cmcx import domains
from cmcx.solver import make_solver
from gpu_util import (ROUNDOFF_FP32, check_fp32_against_spread, component_errors,
                      format_errors, say)

pytestmark = pytest.mark.gpu

STEPS = 5
G = 4
PLAIN, FUSED = 0, 1


@pytest.fixture(scope="module")
def dom():
    d = domains.ellipsoids(256)
    say(f"[C3] {d.notes}; solid fraction {d.solid_fraction():.4f}")
    return d


def _oracle_run(d, dtype):
    say(f"[C3] oracle {dtype}, {STEPS} steps of {d.shape} (CPU, "
        f"{oracle_lib.load().orc_max_threads()} threads)")
    o = oracle_lib.Oracle3D(d, dtype)
    o.create_segments()
    errs = []
    for i in range(STEPS):
        o.update_boundaries()
        r, e = o.step(d.dt, G, 1, True)
        assert r == 0
        errs.append(e)
    f = o.fields()
    segs = [o.get_segments(a) for a in range(3)]
    o.close()
    return f, errs, segs


@pytest.fixture(scope="module")
def ref64(dom):
    return _oracle_run(dom, "f64")


@pytest.fixture(scope="module")
def ref32(dom):
    return _oracle_run(dom, "f32")


def _gpu_run(d, dtype, path, variants):
    s = make_solver(d, dtype, path=path)
    if variants is not None:
        for a, v in enumerate(variants):
            s.set_tile_variant(a, v)
    errs = []
    for i in range(STEPS):
        s.UpdateBoundaries()
        errs.append(s.TimeStep(d.dt, G, 1, True))
    u, v, w, p, T = s.fields()
    s.close()
    return (u, v, w, T), errs


def test_c3_segments_bit_exact(dom, ref64):
    s = make_solver(dom, "f32", path=FUSED)
    for a in range(3):
        got, want = s.segments(a), ref64[2][a]
        say(f"  axis {a}: {len(want)} segments, S = {int(want[:, 6].sum())}")
        assert got.shape == want.shape
        assert np.array_equal(got, want)
    s.close()


@pytest.mark.parametrize("variants", [(6, 6, 6), None], ids=["v666", "default"])
def test_c3_fp32_fused(dom, ref64, ref32, variants):
    got, errs = _gpu_run(dom, "f32", FUSED, variants)
    p = dom.params()
    e64 = component_errors(got, ref64[0], dom.base_t, p.v_T, ROUNDOFF_FP32)
    e32o = component_errors(ref32[0], ref64[0], dom.base_t, p.v_T, ROUNDOFF_FP32)
    say(f"  C3 fp32 fused {variants}: vs FP64 oracle: {format_errors(e64)}")
    say(f"  C3 FP32 oracle vs FP64 oracle: {format_errors(e32o)}")
    for nm in ("u", "v", "w", "p"):
        assert e64[nm][1] < 1e-5, (nm, e64[nm])
    bad = check_fp32_against_spread(e64, e32o)
    assert not bad, bad
    # divergence errors (FP64 sums of FP32 fields) follow the oracle's
    for e, e_ref in zip(errs, ref64[1]):
        say(f"  divergence error {e:.6e} (FP64 oracle {e_ref:.6e})")
        assert abs(e - e_ref) <= 1e-2 * abs(e_ref) + 1e-12, (e, e_ref)


def test_c3_fp32_plain_bit_exact(dom, ref32):
    got, errs = _gpu_run(dom, "f32", PLAIN, None)
    for nm, a, b in zip("uvwT", got, ref32[0]):
        nd = int(np.count_nonzero(a != b.astype(np.float32)))
        say(f"  C3 fp32 plain {nm}: {nd} nodes differ from the FP32 oracle")
        assert nd == 0, nm


def test_c3_fp64_fused(dom, ref64):
    got, errs = _gpu_run(dom, "f64", FUSED, None)
    e64 = component_errors(got, ref64[0], dom.base_t, dom.params().v_T)
    say(f"  C3 fp64 fused: {format_errors(e64)}")
    for nm, (own, sc) in e64.items():
        assert (own if own is not None else sc) < 1e-11, (nm, own, sc)
    for e, e_ref in zip(errs, ref64[1]):
        assert abs(e - e_ref) <= 1e-10 * abs(e_ref) + 1e-300


def test_c3_configured_run_properties(dom):
    s = make_solver(dom, "f32", path=FUSED)
    for a, v in enumerate((6, 6, 6)):
        s.set_tile_variant(a, v)
    err = s.run(dom.steps, dom.dt, G, 1)
    u, v, w, p, T = s.fields()
    say(f"  C3 {dom.steps} steps fp32: divergence error {err:.4e}")
    assert np.isfinite(err) and err < 0.01
    for nm, a in zip("uvwpT", (u, v, w, p, T)):
        assert np.all(np.isfinite(a)), nm
    t = dom.type
    bnd = (t == domains.NODE_BOUND) | (t == domains.NODE_VALVE)
    nsv = bnd & (dom.bc_vel == domains.BC_NOSLIP)
    nst = bnd & (dom.bc_temp == domains.BC_NOSLIP)
    for nm, a, g in (("u", u, dom.vx), ("v", v, dom.vy), ("w", w, dom.vz)):
        assert np.array_equal(a[nsv], g[nsv].astype(np.float32)), nm
    assert np.array_equal(T[nst], dom.T[nst].astype(np.float32))
    # the flow has moved into the domain from the inlet
    fluid = t == domains.NODE_IN
    assert float(np.abs(u[fluid]).max()) > 0.1
    s.close()
