"""Dense brute force of one 3D ADI time step, written from the model (not from
the oracle's code) for tiny domains.  Test infrastructure only.

Model (reference: src/FluidSolver3D/AdiSolver3D.cpp:306-391, 564-666, 732-852):
sweeps Z (cur -> next), Y (next -> half), X (half -> next), each `L` times,
inside `G` Picard iterations.  Every sweep solves, per segment and field, the
system whose rows are
  first node:  BC0  (free: 2 x0 - x1 = 0; no-slip: x0 = node value)
  inner nodes: (-w/2h - nu/h^2) x[p-1] + (3/dt + 2 nu/h^2) x[p] + (w/2h - nu/h^2) x[p+1]
               = 3/dt * in + source
  last node:   BC1  (free: -x[n-2] + 2 x[n-1] = 0; no-slip: x[n-1] = node value)
with w the Picard layer's velocity component along the sweep, nu = v_vis for
u, v, w and t_vis for T, source = -v_T dT/dd for the component along the
sweep and + t_phi Phi_d for T.  Here each system is assembled as a dense matrix
and solved with numpy.linalg.solve.  The layers other than the state start as
copies of it (the reference leaves them uninitialised; the oracle and the
product define them so).
"""
from __future__ import annotations

import numpy as np


def _segments(t, axis):
    """Maximal NODE_IN runs along `axis`, bracketed by their non-IN neighbours,
    lines in the reference's order (X: y, z; Y: x, z; Z: x, y)."""
    nx, ny, nz = t.shape
    segs = []
    if axis == 0:
        order = [(None, j, k) for j in range(ny) for k in range(nz)]
    elif axis == 1:
        order = [(i, None, k) for i in range(nx) for k in range(nz)]
    else:
        order = [(i, j, None) for i in range(nx) for j in range(ny)]
    n = t.shape[axis]
    for o in order:
        def cell(r):
            c = list(o)
            c[axis] = r
            return tuple(c)
        r = 1
        while r < n:
            if t[cell(r)] == 0:
                s = r
                while r < n and t[cell(r)] == 0:
                    r += 1
                if r < n:
                    segs.append([cell(q) for q in range(s - 1, r + 1)])
            r += 1
    return segs


def _d(f, c, axis, h):
    lo = list(c)
    hi = list(c)
    lo[axis] -= 1
    hi[axis] += 1
    return (f[tuple(hi)] - f[tuple(lo)]) / (2 * h[axis])


def _phi(temp, c, axis, h):
    u, v, w = temp[0], temp[1], temp[2]
    if axis == 0:
        ux, vx, wx = _d(u, c, 0, h), _d(v, c, 0, h), _d(w, c, 0, h)
        return 2 * ux * ux + vx * vx + wx * wx + vx * _d(u, c, 1, h) + wx * _d(u, c, 2, h)
    if axis == 1:
        uy, vy, wy = _d(u, c, 1, h), _d(v, c, 1, h), _d(w, c, 1, h)
        return uy * uy + 2 * vy * vy + wy * wy + uy * _d(v, c, 0, h) + wy * _d(v, c, 2, h)
    uz, vz, wz = _d(u, c, 2, h), _d(v, c, 2, h), _d(w, c, 2, h)
    return uz * uz + vz * vz + 2 * wz * wz + uz * _d(w, c, 0, h) + vz * _d(w, c, 1, h)


def _solve_segment(seg, var, axis, src, temp, g, bcv, bct, p, dt, h):
    n = len(seg)
    A = np.zeros((n, n))
    rhs = np.zeros(n)
    free = (bct if var == 3 else bcv)
    nu = p.t_vis if var == 3 else p.v_vis
    vis = nu / h[axis] ** 2
    if free[seg[0]]:
        A[0, 0], A[0, 1] = 2.0, -1.0
    else:
        A[0, 0] = 1.0
        rhs[0] = g[var][seg[0]]
    if free[seg[-1]]:
        A[-1, -2], A[-1, -1] = -1.0, 2.0
    else:
        A[-1, -1] = 1.0
        rhs[-1] = g[var][seg[-1]]
    for r in range(1, n - 1):
        c = seg[r]
        w = temp[axis][c]
        A[r, r - 1] = -w / (2 * h[axis]) - vis
        A[r, r] = 3 / dt + 2 * vis
        A[r, r + 1] = w / (2 * h[axis]) - vis
        s = 0.0
        if var == axis:
            s = -p.v_T * _d(temp[3], c, axis, h)
        elif var == 3:
            s = p.t_phi * _phi(temp, c, axis, h)
        rhs[r] = src[var][c] * 3 / dt + s
    return np.linalg.solve(A, rhs)


def brute_step(dom, p, dt, G, L, state, gpu_order=False):
    """One UpdateBoundaries + TimeStep from `state` = [u, v, w, T] (flat)."""
    shape = dom.shape
    t = dom.type.reshape(shape)
    bcv = dom.bc_vel.reshape(shape).astype(bool)
    bct = dom.bc_temp.reshape(shape).astype(bool)
    g = [a.reshape(shape) for a in (dom.vx, dom.vy, dom.vz, dom.T)]
    h = (dom.dx, dom.dy, dom.dz)
    bound = (t == 2) | (t == 3)
    cur = [a.reshape(shape).copy() for a in state]
    for k in range(4):
        cur[k][bound] = g[k][bound]            # UpdateBoundaries
    # every layer holds the state at the start (the step set-up also copies
    # BOUND/VALVE values into next); nodes no segment covers keep that value
    nxt = [a.copy() for a in cur]
    half = [a.copy() for a in cur]
    temp = [a.copy() for a in cur]
    inside = t == 0
    segs = {a: _segments(t, a) for a in range(3)}

    def sweep(axis, src, out):
        for it in range(L):
            if not gpu_order:
                for seg in segs[axis]:
                    for var in range(4):
                        x = _solve_segment(seg, var, axis, src, temp, g, bcv, bct, p, dt, h)
                        for q, c in enumerate(seg):
                            out[var][c] = x[q]
                for k in range(4):
                    temp[k][inside] = (temp[k][inside] + out[k][inside]) / 2
            else:
                for var in range(4):
                    for seg in segs[axis]:
                        x = _solve_segment(seg, var, axis, src, temp, g, bcv, bct, p, dt, h)
                        for q, c in enumerate(seg):
                            out[var][c] = x[q]
                    temp[var][inside] = (temp[var][inside] + out[var][inside]) / 2

    for _ in range(G):
        sweep(2, cur, nxt)
        sweep(1, nxt, half)
        sweep(0, half, nxt)
        if not gpu_order:
            for k in range(4):
                temp[k][inside] = (temp[k][inside] + nxt[k][inside]) / 2
    return [a.ravel() for a in nxt]
