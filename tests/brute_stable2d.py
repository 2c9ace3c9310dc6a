"""Dense numpy model of the reference's Stable 2D step (src/FluidSolver2D/
StableSolver2D.cpp:21-234), written from the model rather than from the
oracle's loops: the explicit update as array slices, one relaxation sweep as
a triangular solve of the matrix splitting in the sweep's cell order, the
stopping test as a vectorised fold, the layer copies and merges as masked
array assignments.  Test infrastructure for the oracle pins (small grids)."""
import numpy as np
from scipy.linalg import solve_triangular

IN, OUT, BOUND, VALVE = 0, 1, 2, 3


def sweep_order(t):
    """Boundary cells (BOUND, VALVE), then NODE_IN cells, each in (i, j) order."""
    flat = np.arange(t.size).reshape(t.shape)
    b = flat[(t == BOUND) | (t == VALVE)]
    return np.concatenate([b, flat[t == IN]])


def relaxation_matrix(t, dx, dy):
    """A with q_c <- sum_k A[c, k] q_k - rcp dx2 dy2 div_c for the relaxed cells.
    A boundary cell takes each neighbour that is NODE_IN, else the opposite
    one; cells outside the grid or NODE_OUT hold q = 0 (no entry)."""
    nx, ny = t.shape
    dx2, dy2 = dx * dx, dy * dy
    rcp = 0.5 / (dx2 + dy2)
    A = np.zeros((t.size, t.size))

    def is_in(i, j):
        return 0 <= i < nx and 0 <= j < ny and t[i, j] == IN

    def add(c, i, j, w):
        if 0 <= i < nx and 0 <= j < ny and t[i, j] != OUT:
            A[c, i * ny + j] += w

    for i in range(nx):
        for j in range(ny):
            c = i * ny + j
            if t[i, j] == IN:
                for (a, b), w in (((i - 1, j), dy2), ((i + 1, j), dy2), ((i, j - 1), dx2), ((i, j + 1), dx2)):
                    add(c, a, b, rcp * w)
            elif t[i, j] in (BOUND, VALVE):
                up = (i - 1, j) if is_in(i - 1, j) else (i + 1, j)
                dn = (i + 1, j) if is_in(i + 1, j) else (i - 1, j)
                lf = (i, j - 1) if is_in(i, j - 1) else (i, j + 1)
                rt = (i, j + 1) if is_in(i, j + 1) else (i, j - 1)
                add(c, *up, rcp * dy2)
                add(c, *dn, rcp * dy2)
                add(c, *lf, rcp * dx2)
                add(c, *rt, rcp * dx2)
    return A, rcp * dx2 * dy2


def gs_sweep(A, P, x, rhs):
    """One Gauss-Seidel sweep in order P: (I - L) x_P = U x_P(old) + rhs_P."""
    Ap = A[np.ix_(P, P)]
    L = np.tril(Ap, -1)
    U = np.triu(Ap, 1)
    xp = x[P]
    new = solve_triangular(np.eye(len(P)) - L, U @ xp + rhs[P], lower=True)
    out = x.copy()
    out[P] = new
    return out


def fold_err(qn, qo):
    """err after a sweep: the changes |(qn - qo) / qn| in sweep order folded
    with std::max(cur, err) from 0: a NaN resets the running value."""
    with np.errstate(divide="ignore", invalid="ignore"):
        ce = np.abs((qn - qo) / qn)
    nan = np.nonzero(np.isnan(ce))[0]
    if len(nan) == 0:
        return float(max(0.0, ce.max())) if len(ce) else 0.0
    last = nan[-1]
    if last == len(ce) - 1:
        return float("nan")
    return float(ce[last + 1:].max())


def relax(t, dx, dy, div, tol=1e-2, cap=10 ** 7):
    """q of the relaxation from q = 0 and the sweep count (capped at cap)."""
    A, c = relaxation_matrix(t, dx, dy)
    P = sweep_order(t)
    rhs = -c * div.ravel()
    x = np.zeros(t.size)
    n = 0
    while True:
        xn = gs_sweep(A, P, x, rhs)
        err = fold_err(xn[P], x[P])
        x = xn
        n += 1
        if not (err >= tol) or n >= cap:
            return x.reshape(t.shape), n


def central(f, h, axis):
    g = np.zeros_like(f)
    if axis == 0:
        g[1:-1, :] = (f[2:, :] - f[:-2, :]) / (2 * h)
    else:
        g[:, 1:-1] = (f[:, 2:] - f[:, :-2]) / (2 * h)
    return g


def second(f, h, axis):
    g = np.zeros_like(f)
    if axis == 0:
        g[1:-1, :] = (f[2:, :] - 2 * f[1:-1, :] + f[:-2, :]) / (h * h)
    else:
        g[:, 1:-1] = (f[:, 2:] - 2 * f[:, 1:-1] + f[:, :-2]) / (h * h)
    return g


class StableModel:
    """Layers cur, next, temp, next_w as dicts of (nx, ny) arrays; copies and
    merges touch i < nx-1, j < ny-1 only (TimeLayer2D.h:104-172)."""

    def __init__(self, t, gv, dx, dy, startT, v_vis):
        self.t, self.dx, self.dy, self.startT, self.v_vis = t, dx, dy, startT, v_vis
        self.gv = {k: np.array(v, dtype=np.float64) for k, v in gv.items()}
        z = lambda: {k: np.zeros(t.shape) for k in "uvT"}
        self.cur, self.next, self.temp, self.nw = {k: self.gv[k].copy() for k in "uvT"}, z(), z(), z()
        self.copy(self.cur, self.next)
        self.copy(self.cur, self.temp)
        self.sweeps = []

    def inner_mask(self, m):
        k = np.zeros_like(m)
        k[:-1, :-1] = m[:-1, :-1]
        return k

    def copy(self, src, dst, types=(IN, OUT, BOUND, VALVE)):
        m = self.inner_mask(np.isin(self.t, types))
        for k in "uvT":
            dst[k][m] = src[k][m]

    def div_error(self, L):
        t = self.t
        m = (t[:-1, :-1] == IN) & (t[1:, :-1] == IN) & (t[:-1, 1:] == IN) & (t[1:, 1:] == IN)
        U, V = L["u"], L["v"]
        tx = self.dy * (U[1:, :-1] - U[:-1, :-1]) + (U[1:, 1:] - U[:-1, 1:]) / 2
        ty = self.dx * (V[:-1, 1:] - V[:-1, :-1]) + (V[1:, 1:] - V[1:, :-1]) / 2
        e = np.abs(tx + ty)[m]
        return float(e.sum() / len(e)) if len(e) else float("nan")

    def update_boundaries(self):
        m = (self.t == BOUND) | (self.t == VALVE)
        for k in "uvT":
            self.cur[k][m] = self.gv[k][m]
        self.copy(self.cur, self.next, (BOUND, VALVE))

    def step(self, dt, G):
        t, dx, dy = self.t, self.dx, self.dy
        inn = t == IN
        self.copy(self.cur, self.temp)
        self.sweeps = []
        err = self.div_error(self.next)
        it = 0
        while it < G or err > 0.1:
            self.copy(self.cur, self.nw)
            U, V = self.temp["u"], self.temp["v"]
            adv_u = -U * central(U, dx, 0) - V * central(U, dy, 1) + self.v_vis * (second(U, dx, 0) + second(U, dy, 1))
            adv_v = -U * central(V, dx, 0) - V * central(V, dy, 1) + self.v_vis * (second(V, dx, 0) + second(V, dy, 1))
            self.nw["u"][inn] = (self.cur["u"] + dt * adv_u)[inn]
            self.nw["v"][inn] = (self.cur["v"] + dt * adv_v)[inn]
            div = np.where(inn, central(self.nw["u"], dx, 0) + central(self.nw["v"], dy, 1), 0.0)
            q, n = relax(t, dx, dy, div)
            self.sweeps.append(n)
            self.next["u"][inn] = (self.nw["u"] - central(q, dx, 0))[inn]
            self.next["v"][inn] = (self.nw["v"] - central(q, dy, 1))[inn]
            err = self.div_error(self.next)
            m = self.inner_mask(inn)
            for k in "uvT":
                self.temp[k][m] = (self.temp[k][m] + self.next[k][m]) / 2
            it += 1
        out = t == OUT
        self.next["u"][out] = 0.0
        self.next["v"][out] = 0.0
        self.next["T"][out] = self.startT
        self.copy(self.next, self.cur)
        return err, it
