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


class ChannelProblem:
    """The inlet-control gradient workload: a channel with a seeded z-axis
    cylinder (domains.cylinder_channel), loss = mean kinetic energy over a
    `probe`^3 box behind the cylinder (x in [nx/2+1, nx/2+1+probe), centred in
    y and z), controls = the coarse inlet profile (each value fills a 3x3 block
    of inlet nodes) and the kinematic viscosity nu = v_vis."""

    def __init__(self, n: int = 96, steps: int = 200, probe: int = 16, **kw):
        self.d, self.control = domains.cylinder_channel(n=n, steps=steps, **kw)
        d = self.d
        self.steps = steps
        nx, ny, nz = d.shape
        m = np.zeros(d.shape, bool)
        x0 = nx // 2 + 1
        m[x0:x0 + probe, ny // 2 - probe // 2:ny // 2 + probe // 2,
          nz // 2 - probe // 2:nz // 2 + probe // 2] = True
        self.mask = m.ravel()
        self.ids = np.arange(ny * nz)  # inlet face x = 0, (y, z) order
        self.nc = self.control.shape[0]

    def inlet(self, c):
        """Node u values with the inlet face set from control c."""
        d = self.d
        vx = d.vx.copy()
        vx[self.ids] = np.kron(c, np.ones((3, 3)))[:d.ny, :d.nz].ravel()
        return vx

    def block_sum(self, g):
        """Gradient of the node u values -> gradient of the control."""
        d, nc = self.d, self.nc
        g = np.asarray(g, np.float64)[self.ids].reshape(d.ny, d.nz)
        return g[:3 * nc, :3 * nc].reshape(nc, 3, nc, 3).sum(axis=(1, 3))

    def loss(self, fields):
        return probe_ke(fields, self.mask)
