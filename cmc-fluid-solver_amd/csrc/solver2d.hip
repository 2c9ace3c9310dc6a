// 2D ADI time step (the reference's FluidSolver2D path, CPU-only there).
//
// Semantics of AdiSolver2D (src/FluidSolver2D/AdiSolver2D.cpp:21-323,
// TimeLayer2D.h:88-172, Solver2D.cpp:48-85), kept exactly:
//   * one system per grid line (CreateSegments, :228-277): from the non-OUT
//     node before the first NODE_IN to the one after the last; obstacles inside
//     are solved through (BuildMatrix does not look at node types);
//   * rows 1/dt (not 3/dt), BC0 x0 = value or x0 - x1 = 0 (free), BC1 x = value
//     or x[n-2] - x[n-1] = 0 (:74-116), one BC kind per node for all fields;
//   * a direction copies the global Picard layer into a local one, copies the
//     first local solve into it and merges later ones, then copies the local
//     layer and the result out on NODE_IN (SolveDirection, :203-226);
//   * sweeps X then Y, the global loop while it < G or err > 0.1 with the
//     first iteration copying and later ones merging, and the exits above 100
//     iterations or err > 1.0 (TimeStep, :279-323);
//   * every copy and merge skips the last row and column (TimeLayer2D.h:104-172);
//   * the divergence error formula as written (TimeLayer2D.h:88-102).
// Layers other than the state start at zero; their last row and column are
// never written (the reference reads them from freshly allocated memory).
//
// Kernels: one thread per (line, field) runs the Thomas recurrence of
// Algorithms.h:21-38 in the reference's operation order.  The rows of a
// work-group's lines are built by all its threads at once into LDS, and each
// line's recurrence and back substitution then run from there (k2_sweep_lds;
// lines too long for the LDS keep c', d' in a global scratch, k2_sweep), so a
// sweep costs about one division chain per row of its longest line.  With one
// inner iteration a direction is a copy and the sweep, which writes the
// direction's layers itself.  Copies, merges and the divergence sum are
// elementwise; the first G Picard iterations of a step, which run whatever
// their error, are queued back to back and their errors read once.
#pragma clang fp contract(off)

#include "cmcx_internal.h"

#include <algorithm>
#include <memory>
#include <vector>

namespace cmcx {
namespace {

template <class T>
struct L3 {
    T* f[3];  // u, v, T
};

struct Sys2 {
    int a0, a1, b;  // X lines: (i_start, j, i_end); Y lines: (i, j_start, j_end)
};

struct P2 {
    double dt, dx, dy, v_T, v_vis, t_vis, t_phi;
};

constexpr int kBS2 = 256;

inline unsigned blocks2(int64_t n) { return (unsigned)((n + kBS2 - 1) / kBS2); }

// dst <- src on nodes (i < nx-1, j < ny-1) whose type bit is set in `mask`;
// merge: dst <- (dst + src) / 2 instead.
template <class T>
__global__ void k2_copy(int nx, int ny, const uint8_t* __restrict__ type, int mask, int merge,
                        L3<T> src, L3<T> dst) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nx * ny) return;
    const int i = (int)(q / ny), j = (int)(q % ny);
    if (i >= nx - 1 || j >= ny - 1 || !((mask >> type[q]) & 1)) return;
#pragma unroll
    for (int f = 0; f < 3; f++) dst.f[f][q] = merge ? (dst.f[f][q] + src.f[f][q]) / 2 : src.f[f][q];
}

// Fused copies of the one-iteration path, on nodes (i < nx-1, j < ny-1) as
// k2_copy.  k2_fan: d0..d3 <- src (the step's copies of the state, with the
// first X direction's copy of the Picard layer).  k2_merge_fan: temp <- next
// on NODE_IN (merge: (temp + next) / 2), then tl <- temp everywhere (the next
// X direction's copy of the Picard layer).
template <class T>
__global__ void k2_fan(int nx, int ny, L3<T> src, L3<T> d0, L3<T> d1, L3<T> d2, L3<T> d3) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nx * ny) return;
    const int i = (int)(q / ny), j = (int)(q % ny);
    if (i >= nx - 1 || j >= ny - 1) return;
#pragma unroll
    for (int f = 0; f < 3; f++) {
        const T v = src.f[f][q];
        d0.f[f][q] = v;
        d1.f[f][q] = v;
        d2.f[f][q] = v;
        d3.f[f][q] = v;
    }
}

template <class T>
__global__ void k2_merge_fan(int nx, int ny, const uint8_t* __restrict__ type, int merge,
                             L3<T> next, L3<T> temp, L3<T> tl) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nx * ny) return;
    const int i = (int)(q / ny), j = (int)(q % ny);
    if (i >= nx - 1 || j >= ny - 1) return;
    const bool in = type[q] == NODE_IN;
#pragma unroll
    for (int f = 0; f < 3; f++) {
        T v = temp.f[f][q];
        if (in) {
            v = merge ? (v + next.f[f][q]) / 2 : next.f[f][q];
            temp.f[f][q] = v;
        }
        tl.f[f][q] = v;
    }
}

// Solver2D::UpdateBoundaries: node values into the state on BOUND/VALVE, then
// state -> next on BOUND and VALVE (a copy that skips the last row and column).
template <class T>
__global__ void k2_update_boundaries(int nx, int ny, const uint8_t* __restrict__ type, L3<T> g,
                                     L3<T> cur, L3<T> next) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nx * ny) return;
    const uint8_t t = type[q];
    if (t != NODE_BOUND && t != NODE_VALVE) return;
    const int i = (int)(q / ny), j = (int)(q % ny);
    const bool inner = i < nx - 1 && j < ny - 1;
#pragma unroll
    for (int f = 0; f < 3; f++) {
        cur.f[f][q] = g.f[f][q];
        if (inner) next.f[f][q] = g.f[f][q];
    }
}

// Solver2D::ClearOutterCells (Solver2D.cpp:75-85)
template <class T>
__global__ void k2_clear_outer(int64_t n, const uint8_t* __restrict__ type, L3<T> next, T startT) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n || type[q] != NODE_OUT) return;
    next.f[0][q] = 0;
    next.f[1][q] = 0;
    next.f[2][q] = startT;
}

// Row t of a line system of n rows (BuildMatrix, AdiSolver2D.cpp:118-178, and
// ApplyBC0/1, :74-116): a, b, c, d of field `var` at cell ct.  Row 0 leaves a
// unset (the recurrence does not read it).
template <class T, bool DIRX>
__device__ __forceinline__ void row2(int t, int n, int var, int64_t ct, int64_t sxs,
                                     const uint8_t* __restrict__ bc, const L3<T>& g,
                                     const L3<T>& lc, const L3<T>& tl, const P2& prm, T& a, T& b,
                                     T& c, T& d) {
    const T dt = (T)prm.dt, hx = (T)prm.dx, hy = (T)prm.dy;
    const T v_T = (T)prm.v_T, t_phi = (T)prm.t_phi;
    const T v_vis = (T)prm.v_vis, t_vis = (T)prm.t_vis;
    const T vis = DIRX ? (var == 2 ? t_vis / (hx * hx) : v_vis / (hx * hx))
                       : (var == 2 ? t_vis / (hy * hy) : v_vis / (hy * hy));
    const T h = DIRX ? hx : hy;
    const T* U = tl.f[0];
    const T* V = tl.f[1];
    const T* TT = tl.f[2];
    if (t == 0) {
        // ApplyBC0
        if (bc[ct] == BC_NOSLIP) {
            b = 1; c = 0; d = g.f[var][ct];
        } else {
            b = 1; c = -1; d = 0;
        }
    } else if (t < n - 1) {
        const T w = DIRX ? U[ct] : V[ct];
        a = -w / (2 * h) - vis;
        b = 1 / dt + 2 * vis;
        c = w / (2 * h) - vis;
        const T lcv = lc.f[var][ct];
        if (DIRX) {
            if (var == 0) {
                const T Tx = (TT[ct + sxs] - TT[ct - sxs]) / (2 * hx);
                d = lcv / dt - v_T * Tx;
            } else if (var == 1) {
                d = lcv / dt;
            } else {
                const T ux = (U[ct + sxs] - U[ct - sxs]) / (2 * hx);
                const T vx = (V[ct + sxs] - V[ct - sxs]) / (2 * hx);
                const T uy = (U[ct + 1] - U[ct - 1]) / (2 * hy);
                d = lcv / dt + t_phi * (2 * ux * ux + vx * vx + uy * vx);
            }
        } else {
            if (var == 0) {
                d = lcv / dt;
            } else if (var == 1) {
                const T Ty = (TT[ct + 1] - TT[ct - 1]) / (2 * hy);
                d = lcv / dt - v_T * Ty;
            } else {
                const T vx = (V[ct + sxs] - V[ct - sxs]) / (2 * hx);
                const T uy = (U[ct + 1] - U[ct - 1]) / (2 * hy);
                const T vy = (V[ct + 1] - V[ct - 1]) / (2 * hy);
                d = lcv / dt + t_phi * (uy * uy + 2 * vy * vy + vx * uy);
            }
        }
    } else {
        // ApplyBC1
        c = 0;
        if (bc[ct] == BC_NOSLIP) {
            a = 0; b = 1; d = g.f[var][ct];
        } else {
            a = 1; b = -1; d = 0;
        }
    }
}

// One line system per thread (blockIdx.y = field): BuildMatrix, ApplyBC0/1,
// SolveTridiagonal (Algorithms.h:21-38), UpdateSegment.  c', d' in a global
// scratch; for systems too long for the LDS of k2_sweep_lds.
template <class T, bool DIRX>
__global__ void k2_sweep(int nsys, const Sys2* __restrict__ sys, int nx, int ny,
                         const uint8_t* __restrict__ bc, L3<T> g, L3<T> lc, L3<T> tl, L3<T> out,
                         P2 prm, T* __restrict__ sc, T* __restrict__ sd) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    const int var = blockIdx.y;
    const Sys2 q = sys[s];
    const int n = DIRX ? q.b - q.a0 + 1 : q.b - q.a1 + 1;
    auto cell = [&](int t) -> int64_t {
        return DIRX ? (int64_t)(q.a0 + t) * ny + q.a1 : (int64_t)q.a0 * ny + q.a1 + t;
    };
    const int64_t sxs = ny;  // stride of i
    auto idx = [&](int t) { return ((int64_t)t * 3 + var) * nsys + s; };  // (row, field, line)
    T a = 0, b, c, d;
    row2<T, DIRX>(0, n, var, cell(0), sxs, bc, g, lc, tl, prm, a, b, c, d);
    // forward elimination, rows built on the fly (SolveTridiagonal with c[n-1] = 0)
    T cp = c / b, dp = d / b;
    sc[idx(0)] = cp;
    sd[idx(0)] = dp;
    for (int t = 1; t < n; t++) {
        row2<T, DIRX>(t, n, var, cell(t), sxs, bc, g, lc, tl, prm, a, b, c, d);
        const T den = b - a * cp;
        cp = c / den;
        dp = (d - dp * a) / den;
        sc[idx(t)] = cp;
        sd[idx(t)] = dp;
    }
    T x = dp;
    out.f[var][cell(n - 1)] = x;
    for (int t = n - 2; t >= 0; t--) {
        x = sd[idx(t)] - sc[idx(t)] * x;
        out.f[var][cell(t)] = x;
    }
}

// The same solve with the rows in LDS.  A work-group takes `spw` systems of
// one field (blockIdx.y): all its threads build the rows of all its systems
// at once into LDS ([row][system] of (a, b, c, d), and a flag byte per row),
// then thread s runs system s's recurrence from there, c' and d' overwriting
// c and d, and its back substitution.  The rows' global loads thus go out in
// parallel, and the serial chain holds no memory access: its rows come from
// LDS in register batches of kSweepBatch, each batch read while the one before
// is eliminated (a thread's global loads would instead queue behind its own
// scratch stores: one counter covers both on this architecture).
// mode 0: the solution on every row into `out` (the local result layer);
// mode 1 (a direction with one inner iteration): the solution into `lt` and
// `out` on NODE_IN rows off the grid's last row and column, which is what the
// direction's copies leave there (SolveDirection, AdiSolver2D.cpp:203-226:
// the local layer takes the first solve on NODE_IN, the layer and the result
// are copied out on NODE_IN, both copies skipping the last row and column);
// the other rows' stores go to `dm` (the local result layer, which such a
// direction does not read), so that no store waits on a branch.
constexpr int kSweepLdsThreads = 256;
constexpr int kSweepBatch = 8;

template <class T>
struct alignas(4 * sizeof(T)) Row4 {
    T a, b, c, d;
};

template <class T, bool DIRX>
__global__ void __launch_bounds__(kSweepLdsThreads) k2_sweep_lds(
    int nsys, int spw, int nmax, const Sys2* __restrict__ sys, int nx, int ny,
    const uint8_t* __restrict__ type, const uint8_t* __restrict__ bc, L3<T> g, L3<T> lc,
    L3<T> tl, L3<T> out, L3<T> lt, L3<T> dm, int mode, P2 prm) {
    extern __shared__ char smem2[];
    constexpr int R = kSweepBatch;
    const int rs = spw * nmax;  // rows in LDS
    Row4<T>* rw = reinterpret_cast<Row4<T>*>(smem2);
    uint8_t* sf = reinterpret_cast<uint8_t*>(rw + rs);
    const int var = blockIdx.y;
    const int s0 = blockIdx.x * spw;
    const int nloc = min(spw, nsys - s0);
    const int64_t sxs = ny;
    auto cell = [&](const Sys2& q, int t) -> int64_t {
        return DIRX ? (int64_t)(q.a0 + t) * ny + q.a1 : (int64_t)q.a0 * ny + q.a1 + t;
    };
    // rows: consecutive threads on neighbouring cells (X: the same row of
    // neighbouring lines, Y: neighbouring rows of a line)
    for (int e = threadIdx.x; e < rs; e += blockDim.x) {
        const int sl = DIRX ? e % spw : e / nmax;
        const int t = DIRX ? e / spw : e % nmax;
        if (sl >= nloc) continue;
        const Sys2 q = sys[s0 + sl];
        const int n = DIRX ? q.b - q.a0 + 1 : q.b - q.a1 + 1;
        if (t >= n) continue;
        const int64_t ct = cell(q, t);
        Row4<T> r;
        r.a = 0;
        row2<T, DIRX>(t, n, var, ct, sxs, bc, g, lc, tl, prm, r.a, r.b, r.c, r.d);
        const int k = t * spw + sl;
        rw[k] = r;
        if (mode == 1) {
            const int i = DIRX ? q.a0 + t : q.a0;
            const int j = DIRX ? q.a1 : q.a1 + t;
            sf[k] = type[ct] == NODE_IN && i < nx - 1 && j < ny - 1;
        }
    }
    __syncthreads();
    const int sl = threadIdx.x;
    if (sl >= nloc) return;
    const Sys2 q = sys[s0 + sl];
    const int n = DIRX ? q.b - q.a0 + 1 : q.b - q.a1 + 1;
    Row4<T>* my = rw + sl;  // row t at my[t * spw]
    const uint8_t* mf = sf + sl;

    // forward elimination (SolveTridiagonal with c[n-1] = 0); rows past the
    // end are read clamped to the last row and not used
    T cp, dp;
    {
        const Row4<T> r0 = my[0];
        cp = r0.c / r0.b;
        dp = r0.d / r0.b;
        my[0].c = cp;
        my[0].d = dp;
    }
    auto elim = [&](const Row4<T>& r, int t) {
        const T den = r.b - r.a * cp;
        cp = r.c / den;
        dp = (r.d - dp * r.a) / den;
        my[t * spw].c = cp;
        my[t * spw].d = dp;
    };
    int t = 1;
    Row4<T> cur[R];
#pragma unroll
    for (int k = 0; k < R; k++) cur[k] = my[min(t + k, n - 1) * spw];
    while (t + R <= n) {
        Row4<T> nxt[R];
#pragma unroll
        for (int k = 0; k < R; k++) nxt[k] = my[min(t + R + k, n - 1) * spw];
#pragma unroll
        for (int k = 0; k < R; k++) elim(cur[k], t + k);
#pragma unroll
        for (int k = 0; k < R; k++) cur[k] = nxt[k];
        t += R;
    }
#pragma unroll
    for (int k = 0; k < R; k++)
        if (t + k < n) elim(cur[k], t + k);

    // back substitution, rows n-2 down to 0 in batches read one batch ahead
    // (the flag bytes are read in mode 0 too, unused: no load under a branch)
    T* const o1 = out.f[var];
    T* const o2 = mode == 1 ? lt.f[var] : out.f[var];
    T* const od = dm.f[var];
    auto put = [&](int t, T v, uint8_t f) {
        const int64_t ct = cell(q, t);
        const bool w = mode == 0 || f != 0;
        (w ? o1 : od)[ct] = v;
        (w ? o2 : od)[ct] = v;
    };
    T x = dp;
    put(n - 1, x, mf[(n - 1) * spw]);
    T bc_[R], bd_[R];
    uint8_t bf_[R];
    t = n - 2;
#pragma unroll
    for (int k = 0; k < R; k++) {
        const int u = max(t - k, 0) * spw;
        bc_[k] = my[u].c;
        bd_[k] = my[u].d;
        bf_[k] = mf[u];
    }
    while (t - R + 1 >= 0) {
        T nc[R], nd[R];
        uint8_t nf[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int u = max(t - R - k, 0) * spw;
            nc[k] = my[u].c;
            nd[k] = my[u].d;
            nf[k] = mf[u];
        }
#pragma unroll
        for (int k = 0; k < R; k++) {
            x = bd_[k] - bc_[k] * x;
            put(t - k, x, bf_[k]);
        }
#pragma unroll
        for (int k = 0; k < R; k++) {
            bc_[k] = nc[k];
            bd_[k] = nd[k];
            bf_[k] = nf[k];
        }
        t -= R;
    }
#pragma unroll
    for (int k = 0; k < R; k++)
        if (t - k >= 0) {
            x = bd_[k] - bc_[k] * x;
            put(t - k, x, bf_[k]);
        }
}

// TimeLayer2D::EvalDivError: per-block partial sums of |tx + ty| in the field
// type and counts; the finish kernel adds them in block order.
constexpr int kDiv2Blocks = 256;

template <class T>
__global__ void k2_div(int nx, int ny, const uint8_t* __restrict__ type, L3<T> L, T dx, T dy,
                       double* __restrict__ part) {
    __shared__ double ssum[kBS2], scnt[kBS2];
    double sum = 0, cnt = 0;
    const int64_t n = (int64_t)(nx - 1) * (ny - 1);
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q / (ny - 1)), j = (int)(q % (ny - 1));
        const int64_t c = (int64_t)i * ny + j;
        if (type[c] == NODE_IN && type[c + ny] == NODE_IN && type[c + 1] == NODE_IN &&
            type[c + ny + 1] == NODE_IN) {
            const T* U = L.f[0];
            const T* V = L.f[1];
            const T tx = dy * (U[c + ny] - U[c]) + (U[c + ny + 1] - U[c + 1]) / 2;
            const T ty = dx * (V[c + 1] - V[c]) + (V[c + ny + 1] - V[c + ny]) / 2;
            T e = tx + ty;
            sum += (double)(e < 0 ? -e : e);
            cnt += 1;
        }
    }
    ssum[threadIdx.x] = sum;
    scnt[threadIdx.x] = cnt;
    __syncthreads();
    for (int o = kBS2 / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            ssum[threadIdx.x] += ssum[threadIdx.x + o];
            scnt[threadIdx.x] += scnt[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = ssum[0];
        part[2 * blockIdx.x + 1] = scnt[0];
    }
}

// One work-group of kDiv2Blocks threads: the partials come into LDS in one
// parallel load, then thread 0 adds them in block order.
__global__ void __launch_bounds__(kDiv2Blocks) k2_div_finish(const double* __restrict__ part,
                                                             int nb, double* __restrict__ out) {
    __shared__ double sp[2 * kDiv2Blocks];
    for (int k = threadIdx.x; k < 2 * nb; k += blockDim.x) sp[k] = part[k];
    __syncthreads();
    if (threadIdx.x != 0) return;
    double s = 0, c = 0;
    for (int b = 0; b < nb; b++) {
        s += sp[2 * b];
        c += sp[2 * b + 1];
    }
    out[0] = s;
    out[1] = c;
}


// ---- StableSolver2D (StableSolver2D.cpp:21-234): explicit advection-diffusion,
// then a projection whose Poisson equation is relaxed cell by cell in place.

// SolveU / SolveV (StableSolver2D.cpp:55-81) on NODE_IN cells: the state plus
// dt times the advection-diffusion of the Picard layer `tl`, into w.
template <class T>
__global__ void k2s_solve_uv(int nx, int ny, const uint8_t* __restrict__ type, L3<T> cur, L3<T> tl,
                             L3<T> w, T dt, T dx, T dy, T v_vis) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nx * ny || type[q] != NODE_IN) return;
    const int64_t e = ny;
    const T* U = tl.f[0];
    const T* V = tl.f[1];
    const T u = U[q], v = V[q];
    const T ux = (U[q + e] - U[q - e]) / (2 * dx), uy = (U[q + 1] - U[q - 1]) / (2 * dy);
    const T uxx = (U[q + e] - 2 * U[q] + U[q - e]) / (dx * dx);
    const T uyy = (U[q + 1] - 2 * U[q] + U[q - 1]) / (dy * dy);
    const T vx = (V[q + e] - V[q - e]) / (2 * dx), vy = (V[q + 1] - V[q - 1]) / (2 * dy);
    const T vxx = (V[q + e] - 2 * V[q] + V[q - e]) / (dx * dx);
    const T vyy = (V[q + 1] - 2 * V[q] + V[q - 1]) / (dy * dy);
    w.f[0][q] = cur.f[0][q] + dt * (-u * ux - v * uy + v_vis * (uxx + uyy));
    w.f[1][q] = cur.f[1][q] + dt * (-u * vx - v * vy + v_vis * (vxx + vyy));
}

// Project's divergence field (StableSolver2D.cpp:85-97): w.Ux + w.Vy on NODE_IN
// cells, zero elsewhere, stored diagonal by diagonal (cell (i, j) at
// doff[i + j] + i - max(0, i + j - ny + 1)): the relaxation reads the cells of
// one anti-diagonal together, which then lie side by side.
template <class T>
__global__ void k2s_div(int nx, int ny, const uint8_t* __restrict__ type, L3<T> w, T dx, T dy,
                        const int32_t* __restrict__ doff, T* __restrict__ dvd) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nx * ny) return;
    const int i = (int)(q / ny), j = (int)(q % ny), d = i + j;
    const int64_t at = doff[d] + i - max(0, d - ny + 1);
    if (type[q] != NODE_IN) {
        dvd[at] = 0;
        return;
    }
    const T* U = w.f[0];
    const T* V = w.f[1];
    dvd[at] = (U[q + ny] - U[q - ny]) / (2 * dx) + (V[q + 1] - V[q - 1]) / (2 * dy);
}

// Project's last loop (StableSolver2D.cpp:153-164): w minus the gradient of q
// (ScalarField2D::grad, central differences in the field type) on NODE_IN cells.
template <class T>
__global__ void k2s_grad(int nx, int ny, const uint8_t* __restrict__ type, L3<T> w,
                         const T* __restrict__ qf, T dx, T dy, L3<T> proj) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nx * ny || type[q] != NODE_IN) return;
    const T gx = (qf[q + ny] - qf[q - ny]) / (2 * dx);
    const T gy = (qf[q + 1] - qf[q - 1]) / (2 * dy);
    proj.f[0][q] = w.f[0][q] - gx;
    proj.f[1][q] = w.f[1][q] - gy;
}

// The relaxation's stopping test folds the cells' relative changes
// |(qn - qo) / qn| in sweep order with std::max(cur, err) from err = 0, then
// continues while err >= tol: a NaN change (0/0) makes err NaN and the next
// cell's change replaces it, so the test depends only on the changes after
// the last NaN (none of them >= tol, or a NaN last, ends the loop).  A
// sequence's summary: kind 0 (empty), 1 (no NaN), 2 (a NaN, then ge counts
// only the changes after the last one); ge: some change >= tol.
struct ErrFold {
    int kind;
    int ge;
};

// One cell's change into the summary.  |qn - qo| against tol |qn| decides
// without dividing away from the threshold (the margin covers the rounding of
// both sides); close to it, and for zero or non-finite values, the change is
// divided out as the reference does, so the decision is the reference's.
__device__ inline void fold_cell(ErrFold& s, double qn, double qo, double tol) {
    const double d = qn - qo;
    const double a = fabs(d), m = fabs(qn), b = tol * m;
    // infinite or NaN values fail both tests below or decide them as the
    // division would (an infinite change is >= tol), so only tiny |qn| and
    // tiny tolerances need the division's exact rounding
    const bool plain = m >= 1e-290 && (b > 1e-300 || tol == 0.0);
    bool nan = false, ge;
    if (plain && a < b * (1.0 - 1e-12)) ge = false;
    else if (plain && a > b * (1.0 + 1e-12)) ge = true;
    else {
        const double ce = fabs(d / qn);
        nan = ce != ce;
        ge = ce >= tol;
    }
    if (nan) {
        s.kind = 2;
        s.ge = 0;
    } else {
        if (s.kind == 0) s.kind = 1;
        s.ge |= ge ? 1 : 0;
    }
}

// sequence a followed by sequence b
__device__ inline ErrFold fold_cat(ErrFold a, ErrFold b) {
    if (b.kind == 0) return a;
    if (a.kind == 0 || b.kind == 2) return b;
    return ErrFold{a.kind, a.ge | b.ge};
}

// Ordered fold over the work-group (thread order), result in thread 0.
__device__ ErrFold fold_block(ErrFold s, ErrFold* sh) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        ErrFold o;
        o.kind = __shfl_down(s.kind, off);
        o.ge = __shfl_down(s.ge, off);
        if (lane + off < 64) s = fold_cat(s, o);
    }
    if (lane == 0) sh[wave] = s;
    __syncthreads();
    ErrFold r{0, 0};
    if (threadIdx.x == 0)
        for (int k = 0; k < nw; k++) r = fold_cat(r, sh[k]);
    __syncthreads();
    return r;
}

// Cell flags of the relaxation: kind in bits 0-1 (0 outside, 1 NODE_IN, 2
// NODE_BOUND or NODE_VALVE), then whether the neighbour (i-1, j), (i+1, j),
// (i, j-1), (i, j+1) is NODE_IN (bits 2-5; cells beyond the grid are not).
enum { PF_IN = 1, PF_BOUND = 2, PF_UP = 4, PF_DOWN = 8, PF_LEFT = 16, PF_RIGHT = 32 };

// Project's relaxation (StableSolver2D.cpp:99-151): sweeps over the boundary
// cells, then the inner cells, each list in (i, j) order, each cell's new value
// from its neighbours' current ones (those earlier in the order already
// relaxed in this sweep), until the stopping test passes.
//
// One work-group.  A sweep's step k relaxes the boundary cell (r, k - r) and
// the inner cell (r, k - 2 - r) of every row r: every value a cell reads is
// then the one the sequential order gives it (boundary cells read inner
// neighbours from the previous sweep and boundary neighbours from this one;
// inner cells read the boundary cells of this sweep, inner neighbours before
// them from this sweep and after them from the previous one), and the cells of
// one step touch no cell another cell of the step reads, so one barrier per
// step orders everything.
//
// Sweeps run in groups of up to D, member s starting LAG = 4 steps after
// member s - 1 in the same q: member s reads diagonals k - 3 .. k + 1 around
// its step k while member s - 1 writes k + 2 and k + 4 and member s + 1 writes
// k - 4 and k - 6, and every value a member reads from the previous sweep is
// final (its inner front passed it at least one step before), so the group
// gives exactly the values of its sweeps run one after the other.  The
// work-group holds D slices of rp threads (rp >= nx, a multiple of 64):
// thread r of a slice relaxes row r's boundary and inner cells for its member
// (a wave skips the boundary part when none of its rows has a boundary cell
// at the step); the members' waves share each SIMD and hide one another's
// latencies.  The stopping test of each
// member is known when the group ends: if member s passes it first, q goes
// back to the snapshot taken at the group's start (in global memory) and a
// group of s + 1 members runs again, which ends on member s's result.  The
// fold runs per row in the row's own order, then over a slice's threads in
// order: boundary rows, then inner rows.
//
// q lives per row over the row's span [lo, hi] of non-outside cells plus a
// zero guard cell at each end (row r at spans[3r]: column j at
// spans[3r] + 1 + j - lo); outside cells keep q = 0, and so do cells beyond the
// grid (the reference reads outside its arrays there, Grid2D.cpp:50-53 has no
// bounds check).  LDS holds q and the flags (QLDS) or they stay in global
// memory.  At most max_sweeps sweeps per launch; status = {sweeps so far,
// converged}; a launch with resume = 1 continues from qspan.
constexpr int kPoisLag = 4;

template <class T, bool QLDS>
__global__ void __launch_bounds__(1024) k2s_poisson(
    int nx, int ny, int S, int rp, int D, const int32_t* __restrict__ spans,
    const uint8_t* __restrict__ fspan, const int32_t* __restrict__ doff, const T* __restrict__ dvd,
    T* __restrict__ qspan, T* __restrict__ save, double dx2, double dy2, double rcp, double tol,
    int max_sweeps, int resume, int32_t* __restrict__ status) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ ErrFold sh[32];
    __shared__ int s_first;
    T* qs = QLDS ? reinterpret_cast<T*>(smem) : qspan;
    const uint8_t* fl = QLDS ? smem + (size_t)S * sizeof(T) : fspan;
    if (QLDS) {
        for (int c = threadIdx.x; c < S; c += blockDim.x) {
            qs[c] = resume ? qspan[c] : (T)0;
            smem[(size_t)S * sizeof(T) + c] = fspan[c];
        }
    } else if (!resume) {
        for (int c = threadIdx.x; c < S; c += blockDim.x) qs[c] = (T)0;
    }
    __syncthreads();
    const int mem = (int)threadIdx.x / rp;  // member slice, uniform per wave
    const int r = (int)threadIdx.x - mem * rp;
    const bool own = r < nx;
    // padded spans of rows r - 1, r, r + 1: storage [a, b] holds columns
    // lo - 1 .. hi + 1 (column j at base + j); a cell is inside for lo <= j <= hi
    int base[3], lo[3], hi[3], a[3], b[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int rd = r + d - 1;
        const bool ok = own && rd >= 0 && rd < nx;
        a[d] = ok ? spans[3 * rd] : 0;
        lo[d] = ok ? spans[3 * rd + 1] : 0;
        hi[d] = ok ? spans[3 * rd + 2] : -1;
        base[d] = a[d] + 1 - lo[d];
        // a row beyond the grid reads element 0, the first row's zero guard
        b[d] = ok ? a[d] + (hi[d] - lo[d] + 1) + 1 : 0;
    }
    auto clampi = [](int x, int lo_, int hi_) { return min(max(x, lo_), hi_); };
    auto inside = [&](int d, int j) -> bool { return j >= lo[d] && j <= hi[d]; };
    // A neighbour row's column j outside its span clamps onto one of that
    // row's zero guard cells (rows beyond the grid onto element 0, the first
    // row's guard), so the read gives the reference's q = 0 there unmasked.
    // The step's barrier: with q in LDS nothing global is shared between
    // threads within a group, so it waits for the LDS traffic only and the div
    // loads stay in flight across it.
    auto step_barrier = [&]() {
        if (QLDS) asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        else __syncthreads();
    };
    // div of (r, j), diagonal-major: the lanes of a wave read one diagonal's
    // consecutive cells
    auto dv_at = [&](int j) -> double {
        const bool ok = own && j >= 0 && j < ny;
        const int d = ok ? r + j : 0;  // every index stays inside its array
        const int at = ok ? doff[d] + r - max(0, d - ny + 1) : 0;
        return ok ? (double)dvd[at] : 0.0;
    };
    const int K = nx + ny + 1;
    const int wpm = rp / 64;  // waves per member slice
    int sweeps = resume ? status[0] : 0;
    int left = max_sweeps;
    bool converged = false;
    int rerun = 0;  // members of a group run again from the snapshot (uniform)
    while (left > 0) {
        const int g = rerun ? rerun : min(D, left);  // members of this group (uniform)
        if (!rerun && g > 1)
            for (int q = threadIdx.x; q < S; q += blockDim.x) save[q] = qs[q];
        ErrFold fb{0, 0}, fi{0, 0};
        const int steps = K + (g - 1) * kPoisLag;
        const int k0 = mem * kPoisLag;  // this member's first step
        double dA = dv_at(0 - k0 - 2 - r), dB = 0.0;
        for (int k2 = 0; k2 < steps; k2 += 2) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int kk = k2 + h;
                if (kk >= steps) break;  // uniform
                double& dcur = h == 0 ? dA : dB;
                double& dnxt = h == 0 ? dB : dA;
                const int k = kk - k0;
                const bool act = mem < g && k >= 0 && k < K;  // uniform per wave
                const int jb = k - r, ji = jb - 2;
                dnxt = dv_at(k + 1 - 2 - r);
                const int own_lo = a[1] + 1, own_hi = max(b[1] - 1, a[1] + 1);
                const int cb = clampi(base[1] + jb, own_lo, own_hi);
                const int ci = clampi(base[1] + ji, own_lo, own_hi);
                const int ub = clampi(base[0] + jb, a[0], b[0]), db = clampi(base[2] + jb, a[2], b[2]);
                const int ui = clampi(base[0] + ji, a[0], b[0]), di = clampi(base[2] + ji, a[2], b[2]);
                T bq = qs[cb], bl = qs[cb - 1], br = qs[cb + 1], bu = qs[ub], bd = qs[db];
                T iq = qs[ci], il = qs[ci - 1], ir = qs[ci + 1], iu = qs[ui], id = qs[di];
                int bf = fl[cb], jf = fl[ci];
                // the empty asm takes every read, so none is sunk into a branch
                asm volatile("" : "+v"(bq), "+v"(bl), "+v"(br), "+v"(bu), "+v"(bd), "+v"(bf));
                asm volatile("" : "+v"(iq), "+v"(il), "+v"(ir), "+v"(iu), "+v"(id), "+v"(jf));
                if (act) {
                    const int f = bf & (inside(1, jb) ? 0xff : 0);
                    if (f & PF_BOUND) {
                        const double up = (double)bu, dn = (double)bd;
                        const double lf_ = (double)bl, rt = (double)br;
                        const double i0 = (f & PF_UP) ? up : dn, i1 = (f & PF_DOWN) ? dn : up;
                        const double j0 = (f & PF_LEFT) ? lf_ : rt, j1 = (f & PF_RIGHT) ? rt : lf_;
                        // div is zero on boundary cells
                        const double qn = rcp * ((i0 + i1) * dy2 + (j0 + j1) * dx2 - 0.0 * dx2 * dy2);
                        qs[cb] = (T)qn;
                        fold_cell(fb, qn, (double)bq, tol);
                    }
                    const int g2 = jf & (inside(1, ji) ? 0xff : 0);
                    if (g2 & PF_IN) {
                        const double up = (double)iu, dn = (double)id;
                        const double qn =
                            rcp * ((up + dn) * dy2 + ((double)il + (double)ir) * dx2 - dcur * dx2 * dy2);
                        qs[ci] = (T)qn;
                        fold_cell(fi, qn, (double)iq, tol);
                    }
                }
                step_barrier();
            }
        }
        // each member's stopping test: an ordered fold over its slice's
        // waves, boundary rows then inner rows
        {
            const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
            ErrFold vb = fb, vi = fi;
            for (int off = 1; off < 64; off <<= 1) {
                ErrFold ob, oi;
                ob.kind = __shfl_down(vb.kind, off);
                ob.ge = __shfl_down(vb.ge, off);
                oi.kind = __shfl_down(vi.kind, off);
                oi.ge = __shfl_down(vi.ge, off);
                if (lane + off < 64) {
                    vb = fold_cat(vb, ob);
                    vi = fold_cat(vi, oi);
                }
            }
            if (lane == 0) {
                sh[wave] = vb;
                sh[16 + wave] = vi;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int first = -1;
            for (int m = 0; m < g && first < 0; m++) {
                ErrFold all{0, 0};
                for (int w = 0; w < wpm; w++) all = fold_cat(all, sh[m * wpm + w]);
                for (int w = 0; w < wpm; w++) all = fold_cat(all, sh[16 + m * wpm + w]);
                if (!(all.kind != 0 && all.ge)) first = m;
            }
            s_first = first;
        }
        __syncthreads();
        const int first = s_first;
        if (rerun || first + 1 == g) {  // a rerun ends on its last member
            sweeps += g;
            converged = true;
            break;
        }
        if (first >= 0) {
            for (int q = threadIdx.x; q < S; q += blockDim.x) qs[q] = save[q];
            __syncthreads();
            rerun = first + 1;
            continue;
        }
        sweeps += g;
        left -= g;
    }
    __syncthreads();
    if (QLDS)
        for (int q = threadIdx.x; q < S; q += blockDim.x) qspan[q] = qs[q];
    if (threadIdx.x == 0) {
        status[0] = sweeps;
        status[1] = converged ? 1 : 0;
    }
}

// q over the whole grid (zero outside the spans) for the gradient.
template <class T>
__global__ void k2s_unspan(int nx, int ny, const int32_t* __restrict__ spans,
                           const T* __restrict__ qspan, T* __restrict__ qf) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nx * ny) return;
    const int i = (int)(q / ny), j = (int)(q % ny);
    const int o = spans[3 * i], l = spans[3 * i + 1], h = spans[3 * i + 2];
    qf[q] = (j >= l && j <= h) ? qspan[o + 1 + j - l] : (T)0;
}
}  // namespace

struct ISolver2 {
    virtual ~ISolver2() {}
    virtual int update_boundaries() = 0;
    virtual int step(double dt, int G, int L, double* err, int* iters) = 0;
    virtual int get_fields(void* u, void* v, void* p, void* T) = 0;
    virtual int set_poisson(double tol, int64_t max_sweeps) = 0;
    virtual int set_grid(const uint8_t* t, const uint8_t* bc, const double* vx, const double* vy,
                         const double* T) = 0;
    virtual int poisson_sweeps(int32_t* out, int max) = 0;
};

template <class T>
class Solver2D final : public ISolver2 {
   public:
    int nx, ny;
    int64_t N;
    hipStream_t st;
    P2 prm{};
    double startT;
    uint8_t *d_type = nullptr, *d_bc = nullptr;
    T* buf = nullptr;  // 7 layers x 3 fields: g, cur, next, half, temp, tl, nl
    L3<T> g{}, cur{}, next{}, half{}, temp{}, tl{}, nl{};
    Sys2 *d_rows = nullptr, *d_cols = nullptr;
    int nrows = 0, ncols = 0;
    int rows_nmax = 0, cols_nmax = 0;  // longest system of each direction
    T* scratch = nullptr;
    int64_t sc_half = 0;
    double* d_part = nullptr;
    double* d_err = nullptr;  // kErrSlots (sum, count) pairs of the divergence error
    double* h_err = nullptr;  // their pinned host copy
    static constexpr int kErrSlots = 64;
    int method = CMCX2D_ADI;
    // Stable solver: next_w, div, q (whole grid and per-row spans), spans and
    // cell flags of the relaxation, its status words
    L3<T> nw{};
    T *d_nw = nullptr, *d_dv = nullptr, *d_qf = nullptr, *d_qspan = nullptr;
    T* d_save = nullptr;  // q at the start of a relaxation group
    int32_t* d_spans = nullptr;
    int32_t* d_doff = nullptr;  // first cell of each anti-diagonal, diagonal-major
    uint8_t* d_fspan = nullptr;
    int32_t* d_pstat = nullptr;
    int S = 0;
    double ptol = 1e-2;            // POISSON_ERR_THRESHOLD (StableSolver2D.h:23)
    int64_t pcap = 1000000;        // the reference has no cap
    std::vector<int32_t> sweeps;   // relaxation sweeps of each project of the last step
    static constexpr int kPoisChunk = 2048;  // sweeps per launch

    Solver2D(const cmcx2d_domain* d, const cmcx_params* p, int meth, hipStream_t stream) {
        st = stream;
        method = meth;
        nx = d->dims[0];
        ny = d->dims[1];
        N = (int64_t)nx * ny;
        prm = {0.0, d->h[0], d->h[1], p->v_T, p->v_vis, p->t_vis, p->t_phi};
        startT = d->start_t;
        CMCX_HIP(hipMalloc((void**)&d_type, N));
        CMCX_HIP(hipMalloc((void**)&d_bc, N));
        CMCX_HIP(hipMalloc((void**)&buf, sizeof(T) * N * 21));
        CMCX_HIP(hipMemsetAsync(buf, 0, sizeof(T) * N * 21, st));
        L3<T>* Ls[7] = {&g, &cur, &next, &half, &temp, &tl, &nl};
        for (int l = 0; l < 7; l++)
            for (int f = 0; f < 3; f++) Ls[l]->f[f] = buf + (int64_t)(l * 3 + f) * N;
        CMCX_HIP(hipMemcpyAsync(d_type, d->type, N, hipMemcpyHostToDevice, st));
        CMCX_HIP(hipMemcpyAsync(d_bc, d->bc, N, hipMemcpyHostToDevice, st));
        {
            std::vector<T> h(N);
            const double* src[3] = {d->v[0], d->v[1], d->T};
            for (int f = 0; f < 3; f++) {
                for (int64_t i = 0; i < N; i++) h[i] = (T)src[f][i];
                CMCX_HIP(hipMemcpyAsync(g.f[f], h.data(), sizeof(T) * N, hipMemcpyHostToDevice, st));
                // AdiSolver2D::Init: the state starts as the node values everywhere
                CMCX_HIP(hipMemcpyAsync(cur.f[f], h.data(), sizeof(T) * N, hipMemcpyHostToDevice, st));
                CMCX_HIP(hipStreamSynchronize(st));
            }
        }
        if (method == CMCX2D_STABLE) build_relaxation(d->type);
        else build_systems(d->type);
        CMCX_HIP(hipMalloc((void**)&d_part, sizeof(double) * 2 * kDiv2Blocks));
        CMCX_HIP(hipMalloc((void**)&d_err, sizeof(double) * 2 * kErrSlots));
        CMCX_HIP(hipHostMalloc((void**)&h_err, sizeof(double) * 2 * kErrSlots, hipHostMallocDefault));
        if (method == CMCX2D_STABLE) {
            // StableSolver2D::Init (StableSolver2D.cpp:44-52): next and temp from
            // the state (copies that skip the last row and column)
            copy(cur, next, kAll, 0);
            copy(cur, temp, kAll, 0);
        }
        CMCX_HIP(hipStreamSynchronize(st));
    }

    // PrepareIndices (StableSolver2D.cpp:167-188) for the fixed mask: per row
    // the span of non-outside cells and the cells' relaxation flags.
    void build_relaxation(const uint8_t* t) {
        auto T_ = [&](int i, int j) -> int {
            return (i < 0 || j < 0 || i >= nx || j >= ny) ? (int)NODE_OUT : t[(int64_t)i * ny + j];
        };
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                if (T_(i, j) == NODE_IN && (i == 0 || j == 0 || i == nx - 1 || j == ny - 1))
                    throw std::string("2D Stable: a NODE_IN cell on the grid's edge (" + std::to_string(i) +
                                      ", " + std::to_string(j) + ")");
        if (nx > 512) throw std::string("2D Stable: grids of more than 512 rows (dimx) are not supported");
        std::vector<int32_t> sp(3 * (size_t)nx);
        std::vector<uint8_t> fl;
        for (int i = 0; i < nx; i++) {
            int lo = 0, hi = -1;
            for (int j = 0; j < ny; j++)
                if (T_(i, j) != NODE_OUT) {
                    if (hi < lo) lo = j;
                    hi = j;
                }
            // storage: a zero guard cell, columns lo .. hi, a zero guard cell
            sp[3 * i] = (int32_t)fl.size();
            sp[3 * i + 1] = lo;
            sp[3 * i + 2] = hi;
            fl.push_back(0);
            for (int j = lo; j <= hi; j++) {
                const int ty = T_(i, j);
                int f = ty == NODE_IN ? PF_IN : (ty == NODE_BOUND || ty == NODE_VALVE) ? PF_BOUND : 0;
                if (T_(i - 1, j) == NODE_IN) f |= PF_UP;
                if (T_(i + 1, j) == NODE_IN) f |= PF_DOWN;
                if (T_(i, j - 1) == NODE_IN) f |= PF_LEFT;
                if (T_(i, j + 1) == NODE_IN) f |= PF_RIGHT;
                fl.push_back((uint8_t)f);
            }
            fl.push_back(0);
        }
        fl.push_back(0);  // one more guard: the last row's reads stay inside
        // spans, flags and q per span (sized by the mask); the rest once
        CMCX_HIP(hipStreamSynchronize(st));
        hipFree(d_fspan);
        hipFree(d_qspan);
        hipFree(d_save);
        d_fspan = nullptr;
        d_qspan = nullptr;
        d_save = nullptr;
        S = (int)fl.size();
        if (!d_spans) {
            std::vector<int32_t> doff(nx + ny);
            for (int d = 0, o = 0; d < nx + ny; d++) {
                doff[d] = o;
                if (d < nx + ny - 1) o += std::min(d, nx - 1) - std::max(0, d - ny + 1) + 1;
            }
            CMCX_HIP(hipMalloc((void**)&d_doff, sizeof(int32_t) * doff.size()));
            CMCX_HIP(hipMemcpy(d_doff, doff.data(), sizeof(int32_t) * doff.size(), hipMemcpyHostToDevice));
            CMCX_HIP(hipMalloc((void**)&d_spans, sizeof(int32_t) * sp.size()));
            CMCX_HIP(hipMalloc((void**)&d_nw, sizeof(T) * N * 2));
            CMCX_HIP(hipMalloc((void**)&d_dv, sizeof(T) * N));
            CMCX_HIP(hipMalloc((void**)&d_qf, sizeof(T) * N));
            CMCX_HIP(hipMalloc((void**)&d_pstat, sizeof(int32_t) * 4));
            CMCX_HIP(hipMemsetAsync(d_nw, 0, sizeof(T) * N * 2, st));
            nw.f[0] = d_nw;
            nw.f[1] = d_nw + N;
            nw.f[2] = temp.f[2];  // never read or written (the step solves u and v only)
        }
        CMCX_HIP(hipMalloc((void**)&d_fspan, std::max(S, 1)));
        CMCX_HIP(hipMalloc((void**)&d_qspan, sizeof(T) * std::max(S, 1)));
        CMCX_HIP(hipMalloc((void**)&d_save, sizeof(T) * std::max(S, 1)));
        CMCX_HIP(hipMemcpyAsync(d_spans, sp.data(), sizeof(int32_t) * sp.size(), hipMemcpyHostToDevice, st));
        if (S) CMCX_HIP(hipMemcpyAsync(d_fspan, fl.data(), S, hipMemcpyHostToDevice, st));
        CMCX_HIP(hipStreamSynchronize(st));
    }

    // Grid2D::Prepare(t) between steps (FluidSolver2D.cpp:131): new cell types,
    // BC kinds and node values; the line systems (ADI) or the relaxation's
    // lists (Stable) follow the new types, as the reference rebuilds them
    // every step.
    int set_grid(const uint8_t* t, const uint8_t* bc, const double* vx, const double* vy,
                 const double* Tt) override {
        if (!t || !bc || !vx || !vy || !Tt) return CMCX_ERR_ARG;
        for (int64_t i = 0; i < N; i++)
            if (t[i] > 3 || bc[i] > 1) {
                set_error("node type or BC kind out of range at node " + std::to_string(i));
                return CMCX_ERR_ARG;
            }
        if (method == CMCX2D_STABLE) build_relaxation(t);
        else {
            CMCX_HIP(hipStreamSynchronize(st));
            hipFree(d_rows);
            hipFree(d_cols);
            hipFree(scratch);
            d_rows = d_cols = nullptr;
            scratch = nullptr;
            build_systems(t);
        }
        CMCX_HIP(hipMemcpyAsync(d_type, t, N, hipMemcpyHostToDevice, st));
        CMCX_HIP(hipMemcpyAsync(d_bc, bc, N, hipMemcpyHostToDevice, st));
        std::vector<T> h(N);
        const double* src[3] = {vx, vy, Tt};
        for (int f = 0; f < 3; f++) {
            for (int64_t i = 0; i < N; i++) h[i] = (T)src[f][i];
            CMCX_HIP(hipMemcpyAsync(g.f[f], h.data(), sizeof(T) * N, hipMemcpyHostToDevice, st));
            CMCX_HIP(hipStreamSynchronize(st));
        }
        return CMCX_OK;
    }

    ~Solver2D() override {
        hipStreamSynchronize(st);
        hipFree(d_type);
        hipFree(d_bc);
        hipFree(buf);
        hipFree(d_rows);
        hipFree(d_cols);
        hipFree(scratch);
        hipFree(d_part);
        hipFree(d_err);
        hipHostFree(h_err);
        hipFree(d_nw);
        hipFree(d_dv);
        hipFree(d_qf);
        hipFree(d_qspan);
        hipFree(d_save);
        hipFree(d_doff);
        hipFree(d_spans);
        hipFree(d_fspan);
        hipFree(d_pstat);
    }

    // AdiSolver2D::CreateSegments (the mask is fixed, so once).
    void build_systems(const uint8_t* t) {
        auto T_ = [&](int i, int j) { return t[(int64_t)i * ny + j]; };
        std::vector<Sys2> rows, cols;
        for (int j = 0; j < ny; j++) {
            int i = 0;
            while (i < nx && T_(i, j) == NODE_OUT) i++;
            while (i + 1 < nx && T_(i + 1, j) != NODE_IN) i++;
            if (i + 1 >= nx) continue;
            const int a = i;
            i = nx - 1;
            while (i >= 0 && T_(i, j) == NODE_OUT) i--;
            while (i - 1 >= 0 && T_(i - 1, j) != NODE_IN) i--;
            rows.push_back({a, j, i});
        }
        for (int i = 0; i < nx; i++) {
            int j = 0;
            while (j < ny && T_(i, j) == NODE_OUT) j++;
            while (j + 1 < ny && T_(i, j + 1) != NODE_IN) j++;
            if (j + 1 >= ny) continue;
            const int a = j;
            j = ny - 1;
            while (j >= 0 && T_(i, j) == NODE_OUT) j--;
            while (j - 1 >= 0 && T_(i, j - 1) != NODE_IN) j--;
            cols.push_back({i, a, j});
        }
        for (const Sys2& s : rows)
            if (s.b - s.a0 + 1 < 2) throw std::string("2D: a line system shorter than 2 nodes");
        for (const Sys2& s : cols)
            if (s.b - s.a1 + 1 < 2) throw std::string("2D: a line system shorter than 2 nodes");
        nrows = (int)rows.size();
        ncols = (int)cols.size();
        rows_nmax = cols_nmax = 0;
        for (const Sys2& s : rows) rows_nmax = std::max(rows_nmax, s.b - s.a0 + 1);
        for (const Sys2& s : cols) cols_nmax = std::max(cols_nmax, s.b - s.a1 + 1);
        if (nrows) {
            CMCX_HIP(hipMalloc((void**)&d_rows, sizeof(Sys2) * nrows));
            CMCX_HIP(hipMemcpy(d_rows, rows.data(), sizeof(Sys2) * nrows, hipMemcpyHostToDevice));
        }
        if (ncols) {
            CMCX_HIP(hipMalloc((void**)&d_cols, sizeof(Sys2) * ncols));
            CMCX_HIP(hipMemcpy(d_cols, cols.data(), sizeof(Sys2) * ncols, hipMemcpyHostToDevice));
        }
        const int64_t nsys = std::max(nrows, ncols);
        sc_half = (int64_t)std::max(nx, ny) * 3 * std::max<int64_t>(nsys, 1);
        CMCX_HIP(hipMalloc((void**)&scratch, sizeof(T) * 2 * sc_half));
    }

    void copy(L3<T> src, L3<T> dst, int mask, int merge) {
        hipLaunchKernelGGL((k2_copy<T>), dim3(blocks2(N)), dim3(kBS2), 0, st, nx, ny, d_type, mask,
                           merge, src, dst);
        CMCX_LAUNCH_CHECK();
    }

    static constexpr int kAll = 0xF, kIn = 1 << NODE_IN;

    // both directions run k2_sweep_lds with one inner iteration
    bool one_pass(int L) const { return L == 1 && nrows && ncols && lds_spw(true) && lds_spw(false); }

    // Systems per work-group of k2_sweep_lds for a direction (0: its rows do
    // not fit the LDS): a power of two, small enough that the launch spreads
    // over about 128 work-groups (the recurrences are serial; more systems per
    // work-group only lengthen its row-building phase), and within 150 KiB.
    int lds_spw(bool dirx) const {
        const int n = dirx ? nrows : ncols, nmax = dirx ? rows_nmax : cols_nmax;
        const int64_t per = (int64_t)nmax * (sizeof(Row4<T>) + 1);
        int spw = 1;
        while (spw < kSweepLdsThreads && (int64_t)(2 * spw) * per <= 150 * 1024 &&
               (int64_t)n * 3 / (2 * spw) >= 128)
            spw *= 2;
        return per <= 150 * 1024 ? spw : 0;
    }

    // mode 1: k2_sweep_lds writes lt and out on the NODE_IN rows (a direction
    // with one inner iteration), else the solution into out
    void sweep_lds(bool dirx, int spw, int mode, L3<T> lc, L3<T> tlc, L3<T> out, L3<T> lt) {
        const int n = dirx ? nrows : ncols, nmax = dirx ? rows_nmax : cols_nmax;
        const size_t lds = (size_t)spw * nmax * (sizeof(Row4<T>) + 1);
        const dim3 grid((unsigned)((n + spw - 1) / spw), 3);
        if (dirx)
            hipLaunchKernelGGL((k2_sweep_lds<T, true>), grid, dim3(kSweepLdsThreads), lds, st, n,
                               spw, nmax, d_rows, nx, ny, d_type, d_bc, g, lc, tlc, out, lt, nl,
                               mode, prm);
        else
            hipLaunchKernelGGL((k2_sweep_lds<T, false>), grid, dim3(kSweepLdsThreads), lds, st, n,
                               spw, nmax, d_cols, nx, ny, d_type, d_bc, g, lc, tlc, out, lt, nl,
                               mode, prm);
        CMCX_LAUNCH_CHECK();
    }

    void sweep(bool dirx, L3<T> lc, L3<T> lt, L3<T> out) {
        const int n = dirx ? nrows : ncols;
        if (!n) return;
        if (const int spw = lds_spw(dirx)) {
            sweep_lds(dirx, spw, 0, lc, lt, out, L3<T>{});
            return;
        }
        // sc/sd indexing uses the launch's system count as the stride
        if (dirx)
            hipLaunchKernelGGL((k2_sweep<T, true>), dim3(blocks2(n), 3), dim3(kBS2), 0, st, n,
                               d_rows, nx, ny, d_bc, g, lc, lt, out, prm, scratch,
                               scratch + sc_half);
        else
            hipLaunchKernelGGL((k2_sweep<T, false>), dim3(blocks2(n), 3), dim3(kBS2), 0, st, n,
                               d_cols, nx, ny, d_bc, g, lc, lt, out, prm, scratch,
                               scratch + sc_half);
        CMCX_LAUNCH_CHECK();
    }

    // AdiSolver2D::SolveDirection; tl_ready: tl already holds lt (a fused copy)
    void direction(bool dirx, int L, L3<T> lc, L3<T> lt, L3<T> out, bool tl_ready = false) {
        if (!(dirx ? nrows : ncols)) return;  // the reference reads list[0] unconditionally
        if (!tl_ready) copy(lt, tl, kAll, 0);
        if (L == 1) {
            if (const int spw = lds_spw(dirx)) {
                // the solve writes what the copies below would leave on NODE_IN
                // (the local layer is read for the stencils, lt and out are not)
                sweep_lds(dirx, spw, 1, lc, tl, out, lt);
                return;
            }
        }
        for (int it = 0; it < L; it++) {
            sweep(dirx, lc, tl, nl);
            copy(nl, tl, kIn, it == 0 ? 0 : 1);
        }
        copy(tl, lt, kIn, 0);
        copy(nl, out, kIn, 0);
    }

    // The divergence error of Lr into slot `slot` (sum and count), no wait.
    void div_enqueue(L3<T> Lr, int slot) {
        hipLaunchKernelGGL((k2_div<T>), dim3(kDiv2Blocks), dim3(kBS2), 0, st, nx, ny, d_type, Lr,
                           (T)prm.dx, (T)prm.dy, d_part);
        CMCX_LAUNCH_CHECK();
        hipLaunchKernelGGL(k2_div_finish, dim3(1), dim3(kDiv2Blocks), 0, st, d_part, kDiv2Blocks,
                           d_err + 2 * slot);
        CMCX_LAUNCH_CHECK();
    }

    // The first n slots into h_err (waits for the stream).
    void read_errors(int n) {
        CMCX_HIP(hipMemcpyAsync(h_err, d_err, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, st));
        CMCX_HIP(hipStreamSynchronize(st));
    }

    double div_error(L3<T> Lr) {
        div_enqueue(Lr, 0);
        read_errors(1);
        return h_err[0] / h_err[1];
    }

    int update_boundaries() override {
        hipLaunchKernelGGL((k2_update_boundaries<T>), dim3(blocks2(N)), dim3(kBS2), 0, st, nx, ny,
                           d_type, g, cur, next);
        CMCX_LAUNCH_CHECK();
        return CMCX_OK;
    }

    int set_poisson(double tol, int64_t max_sweeps) override {
        if (!(tol > 0) || max_sweeps < 1) return CMCX_ERR_ARG;
        ptol = tol;
        pcap = max_sweeps;
        return CMCX_OK;
    }

    int poisson_sweeps(int32_t* out, int max) override {
        for (int k = 0; k < max && k < (int)sweeps.size(); k++) out[k] = sweeps[k];
        return (int)sweeps.size();
    }

    // Sweeps in flight: as many member slices of rp threads as a 1024-thread
    // work-group holds, at most 8.
    template <bool QLDS>
    void launch_poisson_for(int rp, size_t lds, double dx2, double dy2, double rcp, int chunk, int resume) {
        const int D = std::min(8, 1024 / rp);
        hipLaunchKernelGGL((k2s_poisson<T, QLDS>), dim3(1), dim3(rp * D), lds, st, nx, ny, S, rp, D,
                           d_spans, d_fspan, d_doff, d_dv, d_qspan, d_save, dx2, dy2, rcp, ptol, chunk,
                           resume, d_pstat);
        CMCX_LAUNCH_CHECK();
    }

    // StableSolver2D::Project (StableSolver2D.cpp:83-165): w -> next on NODE_IN.
    // Returns false when the relaxation reached the sweep cap.
    bool project() {
        hipLaunchKernelGGL((k2s_div<T>), dim3(blocks2(N)), dim3(kBS2), 0, st, nx, ny, d_type, nw,
                           (T)prm.dx, (T)prm.dy, d_doff, d_dv);
        CMCX_LAUNCH_CHECK();
        const double dx2 = prm.dx * prm.dx, dy2 = prm.dy * prm.dy, rcp = 0.5 / (dx2 + dy2);
        const int rp = std::max(64, (nx + 63) / 64 * 64);  // threads per set
        const size_t lds = (size_t)S * sizeof(T) + (size_t)S;
        const bool in_lds = lds + 2048 <= 160 * 1024;
        int64_t total = 0;
        bool done = false;
        for (int resume = 0; !done && total < pcap; resume = 1) {
            const int chunk = (int)std::min<int64_t>(kPoisChunk, pcap - total);
            if (in_lds) launch_poisson_for<true>(rp, lds, dx2, dy2, rcp, chunk, resume);
            else launch_poisson_for<false>(rp, 0, dx2, dy2, rcp, chunk, resume);
            int32_t h[2];
            CMCX_HIP(hipMemcpyAsync(h, d_pstat, sizeof(h), hipMemcpyDeviceToHost, st));
            CMCX_HIP(hipStreamSynchronize(st));
            total = h[0];
            done = h[1] != 0;
        }
        sweeps.push_back((int32_t)std::min<int64_t>(total, INT32_MAX));
        if (!done) return false;
        hipLaunchKernelGGL((k2s_unspan<T>), dim3(blocks2(N)), dim3(kBS2), 0, st, nx, ny, d_spans, d_qspan,
                           d_qf);
        hipLaunchKernelGGL((k2s_grad<T>), dim3(blocks2(N)), dim3(kBS2), 0, st, nx, ny, d_type, nw, d_qf,
                           (T)prm.dx, (T)prm.dy, next);
        CMCX_LAUNCH_CHECK();
        return true;
    }

    // StableSolver2D::TimeStep (StableSolver2D.cpp:190-234)
    int stable_step(double dt, int G, double* err_out, int* iters) {
        if (G < 0) return CMCX_ERR_ARG;
        prm.dt = dt;
        sweeps.clear();
        copy(cur, temp, kAll, 0);
        double err = div_error(next);
        int it;
        for (it = 0; it < G || err > 0.1; it++) {
            copy(cur, nw, kAll, 0);
            hipLaunchKernelGGL((k2s_solve_uv<T>), dim3(blocks2(N)), dim3(kBS2), 0, st, nx, ny, d_type, cur,
                               temp, nw, (T)dt, (T)prm.dx, (T)prm.dy, (T)prm.v_vis);
            CMCX_LAUNCH_CHECK();
            if (!project()) {
                if (iters) *iters = it + 1;
                set_error("2D Stable: the Poisson relaxation did not converge within " +
                          std::to_string(pcap) + " sweeps");
                return CMCX_ERR_PICARD;
            }
            err = div_error(next);
            copy(next, temp, kIn, 1);
            if (it > 100 || err > 0.1 * 10) {
                if (err_out) *err_out = err;
                if (iters) *iters = it + 1;
                set_error(it > 100 ? "2D: Picard iterations exceeded 100" : "2D: divergence error above 1.0");
                return CMCX_ERR_PICARD;
            }
        }
        hipLaunchKernelGGL((k2_clear_outer<T>), dim3(blocks2(N)), dim3(kBS2), 0, st, N, d_type, next,
                           (T)startT);
        CMCX_LAUNCH_CHECK();
        copy(next, cur, kAll, 0);
        CMCX_HIP(hipStreamSynchronize(st));
        if (err_out) *err_out = err;
        if (iters) *iters = it;
        return CMCX_OK;
    }

    // AdiSolver2D::TimeStep (AdiSolver2D.cpp:279-323)
    int step(double dt, int G, int L, double* err_out, int* iters) override {
        if (method == CMCX2D_STABLE) return stable_step(dt, G, err_out, iters);
        if (G < 0 || L < 1) return CMCX_ERR_ARG;
        prm.dt = dt;
        const bool fused = one_pass(L);
        if (fused) {
            hipLaunchKernelGGL((k2_fan<T>), dim3(blocks2(N)), dim3(kBS2), 0, st, nx, ny, cur, next,
                               half, temp, tl);
            CMCX_LAUNCH_CHECK();
        } else {
            copy(cur, next, kAll, 0);
            copy(cur, half, kAll, 0);
            copy(cur, temp, kAll, 0);
        }
        bool tl_ready = fused;  // tl holds temp (k2_fan, k2_merge_fan)
        // the error before the loop decides only whether a step with G = 0 iterates
        double err = G == 0 ? div_error(next) : 0.0;
        int it = 0;
        // Iterations below G run whatever their error (only an error above 1.0
        // ends them, with CMCX_ERR_PICARD and the state unchanged), so they are
        // queued back to back with their errors in slots that are read once.
        // Iterations queued after an erroring one change only working layers,
        // which the next step's copies overwrite before anything reads them.
        const int batch = std::min(G, (int)kErrSlots);
        for (int k = 0; k < batch; k++) {
            direction(true, L, cur, temp, half, tl_ready);
            direction(false, L, half, temp, next);
            div_enqueue(next, k);
            if (fused && k + 1 < batch) {
                hipLaunchKernelGGL((k2_merge_fan<T>), dim3(blocks2(N)), dim3(kBS2), 0, st, nx, ny,
                                   d_type, k == 0 ? 0 : 1, next, temp, tl);
                CMCX_LAUNCH_CHECK();
                tl_ready = true;
            } else {
                copy(next, temp, kIn, k == 0 ? 0 : 1);
                tl_ready = false;
            }
        }
        if (batch > 0) {
            read_errors(batch);
            for (int k = 0; k < batch; k++) {
                err = h_err[2 * k] / h_err[2 * k + 1];
                if (k > 100 || err > 0.1 * 10) {
                    if (err_out) *err_out = err;
                    if (iters) *iters = k + 1;
                    set_error(k > 100 ? "2D: Picard iterations exceeded 100"
                                      : "2D: divergence error above 1.0");
                    return CMCX_ERR_PICARD;
                }
            }
            it = batch;
        }
        for (; it < G || err > 0.1; it++) {
            direction(true, L, cur, temp, half);
            direction(false, L, half, temp, next);
            err = div_error(next);
            copy(next, temp, kIn, it == 0 ? 0 : 1);
            if (it > 100 || err > 0.1 * 10) {
                if (err_out) *err_out = err;
                if (iters) *iters = it + 1;
                set_error(it > 100 ? "2D: Picard iterations exceeded 100"
                                   : "2D: divergence error above 1.0");
                return CMCX_ERR_PICARD;
            }
        }
        hipLaunchKernelGGL((k2_clear_outer<T>), dim3(blocks2(N)), dim3(kBS2), 0, st, N, d_type,
                           next, (T)startT);
        CMCX_LAUNCH_CHECK();
        copy(next, cur, kAll, 0);
        CMCX_HIP(hipStreamSynchronize(st));
        if (err_out) *err_out = err;
        if (iters) *iters = it;
        return CMCX_OK;
    }

    int get_fields(void* u, void* v, void* p, void* Tt) override {
        void* dst[3] = {u, v, Tt};
        for (int f = 0; f < 3; f++)
            if (dst[f])
                CMCX_HIP(hipMemcpyAsync(dst[f], cur.f[f], sizeof(T) * N, hipMemcpyDeviceToHost, st));
        CMCX_HIP(hipStreamSynchronize(st));
        if (p) {
            std::vector<T> t(N);
            CMCX_HIP(hipMemcpy(t.data(), cur.f[2], sizeof(T) * N, hipMemcpyDeviceToHost));
            T* pp = (T*)p;
            const T vT = (T)prm.v_T;
            for (int64_t i = 0; i < N; i++) pp[i] = vT * t[i];
        }
        return CMCX_OK;
    }
};

}  // namespace cmcx

struct cmcx2d_solver {
    std::unique_ptr<cmcx::ISolver2> impl;
};

#define CMCX2_GUARD(body)                                                                       \
    try {                                                                                       \
        body                                                                                    \
    } catch (const cmcx::HipError& e) {                                                         \
        cmcx::set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where); \
        return CMCX_ERR_HIP;                                                                    \
    } catch (const std::string& e) {                                                            \
        cmcx::set_error(e);                                                                     \
        return CMCX_ERR_ARG;                                                                    \
    } catch (const std::exception& e) {                                                         \
        cmcx::set_error(e.what());                                                              \
        return CMCX_ERR_ARG;                                                                    \
    }

extern "C" {

int cmcx2d_create(const cmcx2d_domain* d, const cmcx_params* p, int dtype, void* stream,
                  cmcx2d_solver** out) {
    return cmcx2d_create_method(d, p, CMCX2D_ADI, dtype, stream, out);
}

int cmcx2d_create_method(const cmcx2d_domain* d, const cmcx_params* p, int method, int dtype,
                         void* stream, cmcx2d_solver** out) {
    if (!d || !p || !out) return CMCX_ERR_ARG;
    *out = nullptr;
    if (dtype != CMCX_F32 && dtype != CMCX_F64) return CMCX_ERR_ARG;
    if (method != CMCX2D_ADI && method != CMCX2D_STABLE) {
        cmcx::set_error("cmcx2d_create_method: method must be CMCX2D_ADI or CMCX2D_STABLE");
        return CMCX_ERR_ARG;
    }
    if (d->dims[0] < 3 || d->dims[1] < 3 || !d->type || !d->bc || !d->v[0] || !d->v[1] || !d->T)
        return CMCX_ERR_ARG;
    const int64_t N = (int64_t)d->dims[0] * d->dims[1];
    for (int64_t i = 0; i < N; i++)
        if (d->type[i] > 3 || d->bc[i] > 1) {
            cmcx::set_error("node type or BC kind out of range at node " + std::to_string(i));
            return CMCX_ERR_ARG;
        }
    CMCX2_GUARD({
        auto* s = new cmcx2d_solver;
        try {
            if (dtype == CMCX_F32)
                s->impl.reset(new cmcx::Solver2D<float>(d, p, method, (hipStream_t)stream));
            else s->impl.reset(new cmcx::Solver2D<double>(d, p, method, (hipStream_t)stream));
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
        return CMCX_OK;
    })
}

int cmcx2d_update_boundaries(cmcx2d_solver* s) {
    if (!s) return CMCX_ERR_ARG;
    CMCX2_GUARD(return s->impl->update_boundaries();)
}

int cmcx2d_step(cmcx2d_solver* s, double dt, int G, int L, double* err, int* iters) {
    if (!s) return CMCX_ERR_ARG;
    CMCX2_GUARD(return s->impl->step(dt, G, L, err, iters);)
}

int cmcx2d_get_fields(cmcx2d_solver* s, void* u, void* v, void* p, void* T) {
    if (!s) return CMCX_ERR_ARG;
    CMCX2_GUARD(return s->impl->get_fields(u, v, p, T);)
}

int cmcx2d_set_grid(cmcx2d_solver* s, const uint8_t* type, const uint8_t* bc, const double* vx,
                    const double* vy, const double* T) {
    if (!s) return CMCX_ERR_ARG;
    CMCX2_GUARD(return s->impl->set_grid(type, bc, vx, vy, T);)
}

int cmcx2d_set_poisson(cmcx2d_solver* s, double tol, int64_t max_sweeps) {
    if (!s) return CMCX_ERR_ARG;
    CMCX2_GUARD(return s->impl->set_poisson(tol, max_sweeps);)
}

int cmcx2d_poisson_sweeps(cmcx2d_solver* s, int32_t* out, int max, int* count) {
    if (!s || !count || (max > 0 && !out)) return CMCX_ERR_ARG;
    CMCX2_GUARD({
        *count = s->impl->poisson_sweeps(out, max);
        return CMCX_OK;
    })
}

int cmcx2d_destroy(cmcx2d_solver* s) {
    if (!s) return CMCX_ERR_ARG;
    CMCX2_GUARD({
        delete s;
        return CMCX_OK;
    })
}

}  // extern "C"
