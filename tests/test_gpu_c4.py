"""The multi-grid workload at its configured size: 64 lid-driven boxes (seed 0)
with edges drawn from {32, 48, 64, 96} per axis and 1-3 spheres each, Re in
[50, 500], dt 1e-3, 4 Picard iterations, all grids advanced together (one
launch per sweep).

Each grid is compared with its own oracle run of the same 3 steps.  Bars:
FP32 within 1e-5 velocity-scaled relative L2 of the FP64 oracle on u, v, w and
p, and each component's own-relative error (u, v, w; T - baseT; p - v_T baseT)
at most twice the FP32 oracle's own error against the FP64 oracle.  Tile
variants: the autotuner's choice on this batch (X and Y: 128-byte tiles with 4
rows per lane; Z: the Picard layer parked in LDS), the defaults, and the
LDS-parked X/Y variant, each with direct launches and with graph replay.
"""
import numpy as np
import pytest

import oracle_lib
from cmcx import domains
from cmcx.solver import make_batch
from gpu_util import (ROUNDOFF_FP32, check_fp32_against_spread, component_errors,
                      format_errors, say)

pytestmark = pytest.mark.gpu

STEPS = 3
G = 4


@pytest.fixture(scope="module")
def doms():
    ds = domains.batch_boxes(count=64, seed=0)
    say(f"[C4] 64 grids, {sum(d.n for d in ds)} cells, shapes "
        f"{sorted(set(d.shape for d in ds))[:6]}...")
    return ds


def _oracle(d, dtype):
    o = oracle_lib.Oracle3D(d, dtype)
    o.create_segments()
    for i in range(STEPS):
        o.update_boundaries()
        r, _ = o.step(d.dt, G, 1, i == STEPS - 1)
        assert r == 0
    f = o.fields()
    o.close()
    return f


@pytest.fixture(scope="module")
def refs(doms):
    say("[C4] oracle runs (FP64 and FP32) of every grid")
    return [(_oracle(d, "f64"), _oracle(d, "f32")) for d in doms]


@pytest.mark.parametrize("graph", [False, True], ids=["direct", "graph"])
@pytest.mark.parametrize("variants", [(3, 3, 2), (0, 0, 0), (6, 6, 6)],
                         ids=["v332", "v000", "v666"])
def test_c4_batch_fp32(doms, refs, variants, graph):
    s = make_batch(doms, "f32")
    for a, v in enumerate(variants):
        s.set_tile_variant(a, v)
    s.set_graph(graph)
    dt = doms[0].dt
    for i in range(STEPS):
        s.UpdateBoundaries()
        s.TimeStep(dt, G, 1, i == STEPS - 1)
    worst = {}
    for k, d in enumerate(doms):
        u, v, w, p, T = s.fields_of(k)
        r64, r32 = refs[k]
        prm = d.params()
        e = component_errors((u, v, w, T), r64, d.base_t, prm.v_T, ROUNDOFF_FP32)
        e32o = component_errors(r32, r64, d.base_t, prm.v_T, ROUNDOFF_FP32)
        for nm in ("u", "v", "w", "p"):
            assert e[nm][1] < 1e-5, (k, d.name, nm, e[nm])
        bad = check_fp32_against_spread(e, e32o)
        assert not bad, (k, d.name, bad, format_errors(e32o))
        for nm, (own, sc) in e.items():
            if own is not None:
                worst[nm] = max(worst.get(nm, 0.0), own)
    say(f"  C4 {variants} graph={graph}: worst own-relative error per component "
        + ", ".join(f"{nm} {x:.2e}" for nm, x in worst.items()))
    s.close()
