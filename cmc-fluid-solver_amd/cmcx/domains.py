"""Synthetic domains: node mask + boundary data for the benchmark workloads.

A domain is what the reference's ``Grid3D`` holds per node (``Node``:
type, bc_vel, bc_temp, v, T -- src/FluidSolver3D/Grid3D.h:73-88), stored as
structure-of-arrays, x-major with z fastest (index ``(i*ny + j)*nz + k``,
src/FluidSolver3D/Grid3D.cpp:242-245).  The reference builds these arrays from
geometry files; the workloads here are synthetic masks with the choices below
fixed and recorded:

* every solid cell, obstacle interiors included, is NODE_BOUND no-slip with
  v = 0 and T = baseT (the reference's ``Build`` sets wall T to baseT,
  Grid3D.cpp:894); the mask has no NODE_OUT cells;
* all six box faces are non-fluid;
* grid spacing h = 1/(dimx-1) on every axis;
* initial state: v = 0, T = baseT in the fluid.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

NODE_IN, NODE_OUT, NODE_BOUND, NODE_VALVE = 0, 1, 2, 3
BC_NOSLIP, BC_FREE = 0, 1

PR = 0.72       # Prandtl number of every reference 3D config
LAMBDA = 1.4    # adiabatic index of every reference 3D config
BASE_T = 1.0    # Config.h default baseT


@dataclasses.dataclass
class Domain3D:
    """Node arrays of one 3D grid (SoA of the reference's ``Node``)."""
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    base_t: float
    type: np.ndarray      # uint8 [N]
    bc_vel: np.ndarray    # uint8 [N]
    bc_temp: np.ndarray   # uint8 [N]
    vx: np.ndarray        # float64 [N]
    vy: np.ndarray
    vz: np.ndarray
    T: np.ndarray
    # flow parameters (FluidParams fields) and run settings
    re: float = 400.0
    dt: float = 5e-4
    steps: int = 100
    num_global: int = 4
    num_local: int = 1
    name: str = ""
    notes: str = ""

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)

    @property
    def n(self) -> int:
        return self.nx * self.ny * self.nz

    def params(self):
        from .params import FluidParams
        return FluidParams.normalized(self.re, PR, LAMBDA)

    def solid_fraction(self) -> float:
        return float(np.mean(self.type != NODE_IN))

    def copy(self) -> "Domain3D":
        d = dataclasses.replace(self)
        for f in ("type", "bc_vel", "bc_temp", "vx", "vy", "vz", "T"):
            setattr(d, f, getattr(self, f).copy())
        return d


@dataclasses.dataclass
class Domain2D:
    nx: int
    ny: int
    dx: float
    dy: float
    base_t: float
    type: np.ndarray      # uint8 [nx*ny], y fastest
    bc: np.ndarray        # uint8 (CondData2D.type)
    vx: np.ndarray
    vy: np.ndarray
    T: np.ndarray
    re: float = 100.0
    dt: float = 1e-3
    steps: int = 200
    num_global: int = 4
    num_local: int = 1
    name: str = ""


def box(nx: int, ny: int, nz: int, base_t: float = BASE_T) -> Domain3D:
    """Closed box: all faces NODE_BOUND no-slip, fluid inside, v = 0, T = baseT."""
    shape = (nx, ny, nz)
    t = np.full(shape, NODE_IN, np.uint8)
    t[0, :, :] = t[-1, :, :] = NODE_BOUND
    t[:, 0, :] = t[:, -1, :] = NODE_BOUND
    t[:, :, 0] = t[:, :, -1] = NODE_BOUND
    h = 1.0 / (nx - 1)
    z = np.zeros(shape)
    return Domain3D(nx, ny, nz, h, h, h, base_t, t.ravel(), np.zeros(t.size, np.uint8),
                    np.zeros(t.size, np.uint8), z.ravel().copy(), z.ravel().copy(),
                    z.ravel().copy(), np.full(t.size, base_t))


def _set_solid(d: Domain3D, mask: np.ndarray) -> None:
    m = mask.ravel()
    d.type[m] = NODE_BOUND
    d.bc_vel[m] = BC_NOSLIP
    d.bc_temp[m] = BC_NOSLIP
    d.vx[m] = 0.0
    d.vy[m] = 0.0
    d.vz[m] = 0.0
    d.T[m] = d.base_t


def _centres(n: int) -> np.ndarray:
    return np.arange(n, dtype=np.float64)


def lid_sphere(n: int = 128, radius: float = 20.0, re: float = 400.0, dt: float = 5e-4,
               steps: int = 100, num_global: int = 4) -> Domain3D:
    """Lid-driven box with one central sphere (128^3, r = 20 cells).

    Sphere centred at ((n-1)/2,)*3, cell-centre test |c - x| < r.  The lid is
    the top face z = n-1 with u = (1, 0, 0); it is assigned last, so the top
    edges move with the lid.  No-slip everywhere else.
    """
    d = box(n, n, n)
    c = (n - 1) / 2.0
    x = _centres(n)
    r2 = (x[:, None, None] - c) ** 2 + (x[None, :, None] - c) ** 2 + (x[None, None, :] - c) ** 2
    _set_solid(d, r2 < radius * radius)
    lid = np.zeros((n, n, n), bool)
    lid[:, :, -1] = True
    d.vx[lid.ravel()] = 1.0
    d.re, d.dt, d.steps, d.num_global = re, dt, steps, num_global
    d.name = f"lid_sphere_{n}"
    d.notes = f"{n}^3 box, sphere r={radius} at centre, lid u=1 on z={n-1}, Re={re}"
    return d


def _rand_rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def ellipsoids(n: int = 256, count: int = 40, seed: int = 0, axis_lo: float = 8.0,
               axis_hi: float = 40.0, re: float = 1000.0, dt: float = 2e-4, steps: int = 50,
               num_global: int = 4) -> Domain3D:
    """256^3 channel with `count` seeded, randomly rotated ellipsoids removed.

    Centres uniform in the box, semi-axes i.i.d. U[axis_lo, axis_hi] cells,
    orientation from a uniform random unit quaternion (numpy default_rng(seed)).
    Inlet x = 0: NODE_VALVE, Dirichlet u = 1.  Outlet x = n-1: NODE_VALVE with
    bc_vel = bc_temp = BC_FREE.  The x faces are assigned after the other
    faces, so they own their edges.  Other faces no-slip.
    """
    rng = np.random.default_rng(seed)
    d = box(n, n, n)
    solid = np.zeros((n, n, n), bool)
    x = _centres(n)
    for _ in range(count):
        c = rng.uniform(0, n - 1, size=3)
        ax = rng.uniform(axis_lo, axis_hi, size=3)
        R = _rand_rotation(rng)
        r = ax.max()
        lo = np.maximum(np.floor(c - r).astype(int), 0)
        hi = np.minimum(np.ceil(c + r).astype(int) + 1, n)
        X = x[lo[0]:hi[0], None, None] - c[0]
        Y = x[None, lo[1]:hi[1], None] - c[1]
        Z = x[None, None, lo[2]:hi[2]] - c[2]
        # body coordinates: p_body = R^T p
        u = R[0, 0] * X + R[1, 0] * Y + R[2, 0] * Z
        v = R[0, 1] * X + R[1, 1] * Y + R[2, 1] * Z
        w = R[0, 2] * X + R[1, 2] * Y + R[2, 2] * Z
        inside = (u / ax[0]) ** 2 + (v / ax[1]) ** 2 + (w / ax[2]) ** 2 < 1.0
        solid[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] |= inside
    _set_solid(d, solid)
    t = d.type.reshape(n, n, n)
    bv = d.bc_vel.reshape(n, n, n)
    bt = d.bc_temp.reshape(n, n, n)
    vx = d.vx.reshape(n, n, n)
    # inlet
    t[0] = NODE_VALVE
    bv[0] = BC_NOSLIP
    bt[0] = BC_NOSLIP
    vx[0] = 1.0
    # outlet
    t[-1] = NODE_VALVE
    bv[-1] = BC_FREE
    bt[-1] = BC_FREE
    vx[-1] = 0.0
    d.re, d.dt, d.steps, d.num_global = re, dt, steps, num_global
    d.name = f"ellipsoids_{n}_s{seed}"
    d.notes = (f"{n}^3, {count} ellipsoids (seed {seed}, semi-axes U[{axis_lo},{axis_hi}]), "
               f"inlet u=1 at x=0, outlet BC_FREE at x={n-1}, Re={re}")
    return d


def batch_boxes(count: int = 64, seed: int = 0, edges=(32, 48, 64, 96), dt: float = 1e-3,
                steps: int = 100, num_global: int = 4) -> list[Domain3D]:
    """Independent lid-driven boxes with 1-3 random spheres each.

    Edges i.i.d. from `edges` per axis; sphere radius U[0.10, 0.30] x smallest
    edge; centres uniform in the interior; Re ~ U[50, 500]; lid u = 1 on the
    top face z = nz-1 (assigned last).
    """
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        nx, ny, nz = (int(v) for v in rng.choice(edges, size=3))
        d = box(nx, ny, nz)
        m = min(nx, ny, nz)
        solid = np.zeros((nx, ny, nz), bool)
        for _ in range(int(rng.integers(1, 4))):
            r = rng.uniform(0.10, 0.30) * m
            c = np.array([rng.uniform(1, nx - 2), rng.uniform(1, ny - 2), rng.uniform(1, nz - 2)])
            r2 = ((_centres(nx)[:, None, None] - c[0]) ** 2 + (_centres(ny)[None, :, None] - c[1]) ** 2
                  + (_centres(nz)[None, None, :] - c[2]) ** 2)
            solid |= r2 < r * r
        _set_solid(d, solid)
        lid = np.zeros((nx, ny, nz), bool)
        lid[:, :, -1] = True
        d.type[lid.ravel()] = NODE_BOUND
        d.vx[lid.ravel()] = 1.0
        d.re = float(rng.uniform(50, 500))
        d.dt, d.steps, d.num_global = dt, steps, num_global
        d.name = f"batch{b}_{nx}x{ny}x{nz}"
        out.append(d)
    return out


def cylinder_channel(n: int = 96, radius: float = 12.0, seed: int = 0, nu: float = 1e-3,
                     control: np.ndarray | None = None, steps: int = 200,
                     dt: float = 3e-3, num_global: int = 4, u0: float = 1.0) -> tuple[Domain3D, np.ndarray]:
    """96^3 channel with a seeded z-axis cylinder; inlet driven by a 32x32 control.

    Cylinder centre (x, y) uniform in the upstream half (x in [r+4, n/2],
    y in [r+4, n-r-5]); inlet x = 0 NODE_VALVE with u = control value repeated
    over 3x3 cell blocks; initial control 1 + 0.05 N(0,1) (default_rng(seed));
    outlet x = n-1 NODE_VALVE BC_FREE; other faces no-slip.  Re = 1/nu.
    Initial state: u = u0 on NODE_IN nodes (an impulsively started cylinder in
    a uniform stream), v = w = 0, T = baseT.  With dt = 3e-3 the 200 steps
    carry the inlet control 0.6 channel lengths downstream (u t = 0.6), past
    the cylinder into the probe behind it; starting from rest, the flow would
    not enter beyond a few cells (the momentum rows carry no pressure
    projection, only viscous and T-gradient coupling).  dt = 5e-3 overflows
    after about 150 steps at the stagnation point in front of the cylinder.
    """
    rng = np.random.default_rng(seed)
    d = box(n, n, n)
    cx = rng.uniform(radius + 4, n / 2)
    cy = rng.uniform(radius + 4, n - radius - 5)
    x = _centres(n)
    r2 = (x[:, None, None] - cx) ** 2 + (x[None, :, None] - cy) ** 2 + 0 * x[None, None, :]
    _set_solid(d, r2 < radius * radius)
    if control is None:
        control = 1.0 + 0.05 * rng.normal(size=(n // 3, n // 3))
    t = d.type.reshape(n, n, n)
    d.vx[d.type == NODE_IN] = u0
    t[0] = NODE_VALVE
    t[-1] = NODE_VALVE
    d.bc_vel.reshape(n, n, n)[-1] = BC_FREE
    d.bc_temp.reshape(n, n, n)[-1] = BC_FREE
    d.vx.reshape(n, n, n)[0] = np.kron(control, np.ones((3, 3)))[:n, :n]
    d.re = 1.0 / nu
    d.dt, d.steps, d.num_global = dt, steps, num_global
    d.name = f"cylinder_channel_{n}"
    return d, control


def channel2d(nx: int = 256, ny: int = 128, radius: float = 16.0, centre=(64, 64),
              u_max: float = 1.0, re: float = 100.0, dt: float = 1e-3, steps: int = 200,
              num_global: int = 4) -> Domain2D:
    """2D channel with a circular obstacle (FP64 CPU workload of the 2D ADI).

    Cells whose centre lies inside the circle are NODE_BOUND no-slip.  Inlet
    x = 0: NODE_VALVE, u = 4 u_max y (1 - y), y = j/(ny-1).  Walls y = 0 and
    y = ny-1: NODE_BOUND no-slip (assigned last, so they own the corners).
    Outlet x = nx-1: NODE_VALVE with BC_FREE (zero gradient in 2D).
    """
    t = np.full((nx, ny), NODE_IN, np.uint8)
    bc = np.zeros((nx, ny), np.uint8)
    vx = np.zeros((nx, ny))
    vy = np.zeros((nx, ny))
    T = np.full((nx, ny), BASE_T)
    ii = _centres(nx)[:, None]
    jj = _centres(ny)[None, :]
    solid = (ii - centre[0]) ** 2 + (jj - centre[1]) ** 2 < radius * radius
    t[solid] = NODE_BOUND
    t[0, :] = NODE_VALVE
    yy = np.arange(ny) / (ny - 1)
    vx[0, :] = 4 * u_max * yy * (1 - yy)
    t[-1, :] = NODE_VALVE
    bc[-1, :] = BC_FREE
    t[:, 0] = NODE_BOUND
    t[:, -1] = NODE_BOUND
    bc[:, 0] = bc[:, -1] = BC_NOSLIP
    vx[:, 0] = vx[:, -1] = 0.0
    h = 1.0 / (nx - 1)
    return Domain2D(nx, ny, h, h, BASE_T, t.ravel(), bc.ravel(), vx.ravel(), vy.ravel(),
                    T.ravel(), re=re, dt=dt, steps=steps, num_global=num_global,
                    name=f"channel2d_{nx}x{ny}")


def random_domain(shape, rng: np.random.Generator, solid_frac: float = 0.15,
                  free_faces: bool = False, lid: bool = True) -> Domain3D:
    """Small random well-formed domain for parity tests (random solid cells)."""
    nx, ny, nz = shape
    d = box(nx, ny, nz)
    inner = rng.random((nx, ny, nz)) < solid_frac
    inner[0] = inner[-1] = False
    _set_solid(d, inner)
    if lid:
        top = np.zeros((nx, ny, nz), bool)
        top[:, :, -1] = True
        d.vx[top.ravel()] = 1.0
        d.vy[top.ravel()] = 0.3
    if free_faces:
        t = d.type.reshape(shape)
        t[-1] = NODE_VALVE
        d.bc_vel.reshape(shape)[-1] = BC_FREE
        d.bc_temp.reshape(shape)[-1] = BC_FREE
        t[0] = NODE_VALVE
        d.vx.reshape(shape)[0] = 0.5
    return d
