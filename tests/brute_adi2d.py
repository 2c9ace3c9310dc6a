"""Dense brute force of one 2D ADI time step, written from the model (not from
the oracle's code) for tiny domains.  Test infrastructure only.

Model (reference: src/FluidSolver2D/AdiSolver2D.cpp:118-323,
TimeLayer2D.h:88-172, Solver2D.cpp:48-85):
  * one segment per grid line, from the first non-OUT node whose successor is
    NODE_IN to the last such node from the other end; obstacles inside are
    solved through;
  * rows p = 1..n-2: (-w/2h - nu/h^2) x[p-1] + (1/dt + 2 nu/h^2) x[p] +
    (w/2h - nu/h^2) x[p+1] = in/dt + source, w the local Picard layer's
    velocity along the sweep, source -v_T dT/dd for the component along the
    sweep and + t_phi Phi_d for T (Phi_X = 2 u_x^2 + v_x^2 + u_y v_x,
    Phi_Y = u_y^2 + 2 v_y^2 + v_x u_y);
  * first row: no-slip x0 = node value, free x0 - x1 = 0; last row: no-slip
    x = node value, free x[n-2] - x[n-1] = 0;
  * a direction starts its local Picard layer from the global one (all node
    types), copies the first local result into it and merges later ones; at
    the end the local layer and the result go to the global layer and the
    output on NODE_IN; all copies and merges skip the last row and column;
  * the global loop runs while it < G or err > 0.1 (err: the divergence
    formula as written, TimeLayer2D.h:96-97), first iteration copies, later
    ones merge; OUT nodes are cleared to (0, 0, startT) at the end.
Each system is assembled densely and solved with numpy.linalg.solve.  Layers
other than the state start at zero (the reference's fresh heap memory).
"""
from __future__ import annotations

import numpy as np

IN, OUT, BOUND, VALVE = 0, 1, 2, 3
FREE = 1


def segments2d(t):
    nx, ny = t.shape
    rows, cols = [], []
    for j in range(ny):  # X-direction systems, one per row
        line = t[:, j]
        a = next((i for i in range(nx - 1) if line[i] != OUT and line[i + 1] == IN), None)
        if a is None:
            continue
        b = next(i for i in range(nx - 1, 0, -1) if line[i] != OUT and line[i - 1] == IN)
        rows.append((a, j, b))
    for i in range(nx):  # Y-direction systems, one per column
        line = t[i, :]
        a = next((j for j in range(ny - 1) if line[j] != OUT and line[j + 1] == IN), None)
        if a is None:
            continue
        b = next(j for j in range(ny - 1, 0, -1) if line[j] != OUT and line[j - 1] == IN)
        cols.append((i, a, b))
    return rows, cols


def _d(f, i, j, di, dj, h):
    return (f[i + di, j + dj] - f[i - di, j - dj]) / (2 * h)


def brute_step2d(dom, p, dt, G, L, state, max_iters=100):
    nx, ny = dom.nx, dom.ny
    t = dom.type.reshape(nx, ny)
    bc = dom.bc.reshape(nx, ny)
    gv = [dom.vx.reshape(nx, ny), dom.vy.reshape(nx, ny), dom.T.reshape(nx, ny)]
    dx, dy = dom.dx, dom.dy
    cur = [np.array(a, np.float64).reshape(nx, ny).copy() for a in state]
    inner = np.zeros((nx, ny), bool)
    inner[:nx - 1, :ny - 1] = True
    inm = inner & (t == IN)
    rows, cols = segments2d(t)

    def copy_inner(src, mask):
        out = [np.zeros((nx, ny)) for _ in range(3)]
        for k in range(3):
            out[k][mask] = src[k][mask]
        return out

    def solve_dir(segs, dirx, lc, temp, out):
        tl = copy_inner(temp, inner)
        nl = [np.zeros((nx, ny)) for _ in range(3)]
        h = dx if dirx else dy
        for it in range(L):
            for (a0, a1, b) in segs:
                cells = [(a0 + q, a1) for q in range(b - a0 + 1)] if dirx else \
                        [(a0, a1 + q) for q in range(b - a1 + 1)]
                n = len(cells)
                for var in range(3):
                    vis = (p.t_vis if var == 2 else p.v_vis) / (h * h)
                    A = np.zeros((n, n))
                    r = np.zeros(n)
                    for q in (0, n - 1):
                        i, j = cells[q]
                        if bc[i, j] == FREE:
                            if q == 0:
                                A[0, 0], A[0, 1] = 1.0, -1.0
                            else:
                                A[n - 1, n - 2], A[n - 1, n - 1] = 1.0, -1.0
                        else:
                            A[q, q] = 1.0
                            r[q] = gv[var][i, j]
                    for q in range(1, n - 1):
                        i, j = cells[q]
                        w = tl[0][i, j] if dirx else tl[1][i, j]
                        A[q, q - 1] = -w / (2 * h) - vis
                        A[q, q] = 1 / dt + 2 * vis
                        A[q, q + 1] = w / (2 * h) - vis
                        r[q] = lc[var][i, j] / dt
                        U, V, T = tl
                        if dirx:
                            ux, vx, uy = _d(U, i, j, 1, 0, dx), _d(V, i, j, 1, 0, dx), _d(U, i, j, 0, 1, dy)
                            if var == 0:
                                r[q] -= p.v_T * _d(T, i, j, 1, 0, dx)
                            elif var == 2:
                                r[q] += p.t_phi * (2 * ux * ux + vx * vx + uy * vx)
                        else:
                            vx, uy, vy = _d(V, i, j, 1, 0, dx), _d(U, i, j, 0, 1, dy), _d(V, i, j, 0, 1, dy)
                            if var == 1:
                                r[q] -= p.v_T * _d(T, i, j, 0, 1, dy)
                            elif var == 2:
                                r[q] += p.t_phi * (uy * uy + 2 * vy * vy + vx * uy)
                    x = np.linalg.solve(A, r)
                    for q, (i, j) in enumerate(cells):
                        nl[var][i, j] = x[q]
            for k in range(3):
                tl[k][inm] = nl[k][inm] if it == 0 else (tl[k][inm] + nl[k][inm]) / 2
        for k in range(3):
            temp[k][inm] = tl[k][inm]
            out[k][inm] = nl[k][inm]

    def div(lay):
        U, V = lay[0], lay[1]
        err, cnt = 0.0, 0
        for i in range(nx - 1):
            for j in range(ny - 1):
                if t[i, j] == IN and t[i + 1, j] == IN and t[i, j + 1] == IN and t[i + 1, j + 1] == IN:
                    tx = dy * (U[i + 1, j] - U[i, j]) + (U[i + 1, j + 1] - U[i, j + 1]) / 2
                    ty = dx * (V[i, j + 1] - V[i, j]) + (V[i + 1, j + 1] - V[i + 1, j]) / 2
                    err += abs(tx + ty)
                    cnt += 1
        return err / cnt

    nxt = copy_inner(cur, inner)
    half = copy_inner(cur, inner)
    temp = copy_inner(cur, inner)
    err = div(nxt)
    it = 0
    status = 0
    while it < G or err > 0.1:
        solve_dir(rows, True, cur, temp, half)
        solve_dir(cols, False, half, temp, nxt)
        err = div(nxt)
        for k in range(3):
            temp[k][inm] = nxt[k][inm] if it == 0 else (temp[k][inm] + nxt[k][inm]) / 2
        if it > max_iters:
            status = 2
            break
        if err > 1.0:
            status = 3
            break
        it += 1
    if status:
        return status, err, it + 1, cur
    outm = t == OUT
    nxt[0][outm] = 0.0
    nxt[1][outm] = 0.0
    nxt[2][outm] = dom.base_t
    for k in range(3):
        cur[k][inner] = nxt[k][inner]
    return 0, err, it, cur
