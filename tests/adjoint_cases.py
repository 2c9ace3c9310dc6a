"""Small domains and losses shared by the adjoint tests (oracle and GPU)."""
from __future__ import annotations

import numpy as np

from cmcx import domains

STEPS, G, DT = 3, 2, 4e-3


def small_domain():
    """9x8x7 box: random solid cells, lid (u=1, v=0.3), a VALVE inlet at x=0
    with non-uniform u, v, T and a BC_FREE VALVE outlet at x=nx-1; Re=50."""
    rng = np.random.default_rng(5)
    d = domains.random_domain((9, 8, 7), rng, 0.08, free_faces=True)
    sh = d.shape
    t = d.type.reshape(sh)
    t[0] = domains.NODE_VALVE
    d.vx.reshape(sh)[0] = 0.5 + rng.random(sh[1:])
    d.vy.reshape(sh)[0] = 0.1 * rng.random(sh[1:])
    d.T.reshape(sh)[0] = 1.0 + 0.2 * rng.random(sh[1:])
    d.re = 50.0
    d.dt = DT
    return d


def probe_mask(shape, size=16, x0=None):
    """Probe box: x in [x0, x0+size) (default 2n/3), centred in y and z."""
    nx, ny, nz = shape
    m = np.zeros(shape, bool)
    if x0 is None:
        x0 = min(2 * nx // 3, nx - 1 - size)
    m[x0:x0 + size, ny // 2 - size // 2:ny // 2 + size // 2, nz // 2 - size // 2:nz // 2 + size // 2] = True
    return m.ravel()


def probe_ke(fields, mask):
    """Mean kinetic energy 1/2 (u^2 + v^2 + w^2) over the probe, and its seed."""
    u, v, w = (np.asarray(fields[k], np.float64) for k in range(3))
    n = mask.sum()
    loss = 0.5 * float(np.sum((u * u + v * v + w * w)[mask])) / n
    seed = np.zeros((4, u.size))
    for k, a in enumerate((u, v, w)):
        seed[k][mask] = a[mask] / n
    return loss, seed
