// Discrete adjoint of the fused ADI sweep (reverse mode): two kernels per
// reversed sweep.
//
// Forward sweep along axis d, per line (block diagonal over the segments):
//   M(temp) x = r(in, temp);  out = x on segment nodes;
//   temp' = alpha_f temp + beta_f x on NODE_IN nodes (the Picard merge: 1/2, 1/2,
//   or 1/4, 3/4 after the X sweep, where the reference merges twice).
// Reverse, given xbar (adjoint of out, NODE_IN nodes) and tbar (adjoint of temp'):
//   adj_solve  (line-wide)  M^T lam = xbar + beta_f tbar on NODE_IN rows, 0 on
//              end rows; lam on Dirichlet end rows is the node-value gradient.
//              Row p of M^T is (c_{p-1}, b_p, a_{p+1}): the coefficients of the
//              neighbouring rows are rebuilt from their roles and temp.
//              xbar is never stored between sweeps: it is the adjoint of the
//              next sweep's input, (3/dt) lam of the sweep reversed before (read
//              from lam, which this solve then overwrites row by row in the
//              same thread), zero after a Z sweep (whose input is the state),
//              or the seed / the previous reversed step's state adjoint.
//   adj_point  (pointwise)  with Mbar = -lam x^T on the tridiagonal pattern:
//              tbar <- alpha tbar (NODE_IN) + d(a, c)/dw : Mbar (advecting
//              velocity) + the transposed pressure and dissipation stencils of
//              the neighbouring rows (a gather: each node collects the terms of
//              the rows that read it, so no two threads write one node);
//              inbar = (3/dt) lam (the adjoint of the sweep input, taken by
//              the next solve from lam, or for a Z sweep, whose input is the
//              state, added to the state adjoint's accumulator);
//              per-block partial sums of the FluidParams gradients.
//   end of a reversed step (folded into the last adj_point of the step):
//              state adjoint = (Z inputs) + tbar on NODE_IN nodes; tbar on
//              BOUND/VALVE nodes goes to the node values (UpdateBoundaries
//              wrote them into the state and the Picard layer).
// Row definitions are those of the forward (AdiSolver3D.cpp:755-852,
// TimeLayer3D.h:338-343, 554-588); the oracle's Adjoint restates the same
// transpose per segment on the CPU.
#include "cmcx_internal.h"
#include "line_solve.h"

#include <algorithm>
#include <atomic>
#include <vector>

namespace cmcx {

namespace {

using namespace lsolve;

constexpr int kAdjChunk = 8;
constexpr int kPointBlocks = 2048;  // fixed grid of adj_point (per-block parameter partials)

template <class T>
struct Coef {
    T av, bv, cv, at, bt, ct;
};

// Coefficients of one row of M (both matrices): interior row from the advection
// term adv = w / 2h, boundary and identity rows from the role and BC kinds (the
// rows of bc_row: a free start (0, 2, -1), a free end (-1, 2, 0), else the
// identity).  Written with selects only: a branch here would pull the loads of
// w and the BC kinds into it, each waited for on its own.
template <class T>
__device__ __forceinline__ Coef<T> row_coef(int role, bool fv, bool ft, T adv,
                                            const SweepParams<T>& p) {
    const bool in = role == RL_IN, st = role == RL_START, en = role == RL_END;
    const bool sv = st & fv, ev = en & fv, s_t = st & ft, e_t = en & ft;
    Coef<T> k;
    k.av = in ? -adv - p.vis_v : (ev ? (T)-1 : (T)0);
    k.cv = in ? adv - p.vis_v : (sv ? (T)-1 : (T)0);
    k.bv = in ? p.three_dt + 2 * p.vis_v : ((sv | ev) ? (T)2 : (T)1);
    k.at = in ? -adv - p.vis_t : (e_t ? (T)-1 : (T)0);
    k.ct = in ? adv - p.vis_t : (s_t ? (T)-1 : (T)0);
    k.bt = in ? p.three_dt + 2 * p.vis_t : ((s_t | e_t) ? (T)2 : (T)1);
    return k;
}

// Transposed line solve.  Work-group = LPG neighbouring lines x all rows, in the
// layout of the strided forward kernel; for the Z axis the lines of a tile are
// LPG consecutive y at one x (rows contiguous in memory).
template <class T, int AXIS, int M, int LPG, int MAXT>
__global__ void __launch_bounds__(MAXT) adj_solve_kernel(
    Geom g, SweepParams<T> p, const uint8_t* __restrict__ type, const uint8_t* __restrict__ bcv,
    const uint8_t* __restrict__ bct, Layer4<T> tin, Layer4<T> xs, T xscale, Layer4<T> tb, T beta,
    Layer4<T> lam, const int32_t* __restrict__ tiles, int ltiles) {
    extern __shared__ char smem_raw[];
    T* red = reinterpret_cast<T*>(smem_raw);
    constexpr int GPW = 64 / LPG;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const int gl = lane % LPG;
    const int chunk = wave * GPW + lane / LPG;
    const int P = nw * GPW;
    const RedLayout L{2 * P, kRedFields * 2 * P + 1};

    const int tile = tiles[xcd_remap(blockIdx.x, gridDim.x)];
    const int outer = tile / ltiles;
    const int line = (tile % ltiles) * LPG + gl;
    int n;
    int64_t s, base;
    bool ok;
    if (AXIS == AX_X) {
        n = g.nx; s = g.sx; base = (int64_t)outer * g.sy + line; ok = line < g.nz;
    } else if (AXIS == AX_Y) {
        n = g.ny; s = g.sy; base = (int64_t)outer * g.sx + line; ok = line < g.nz;
    } else {
        n = g.nz; s = 1; base = (int64_t)outer * g.sx + (int64_t)line * g.sy; ok = line < g.ny;
    }
    const int r0 = chunk * M;

    // Every load of the thread goes out in one batch: node types of rows
    // r0-2 .. r0+M+1, the advecting velocity and BC kinds of rows r0-1 .. r0+M,
    // xbar and tbar of rows r0 .. r0+M-1, from addresses clamped to node 0
    // outside the line and selected afterwards.  A value used under a condition
    // only would be loaded there, behind its own wait: the velocity and the BC
    // kinds are therefore used once, all together, right after the loads.
    uint8_t ty[M + 4];
#pragma unroll
    for (int j = 0; j < M + 4; j++) {
        const int r = r0 + j - 2;
        const bool inr = ok && r >= 0 && r < n;
        const uint8_t t = type[inr ? base + (int64_t)r * s : 0];
        ty[j] = inr ? t : kOutside;
    }
    const T* wv = tin.f[AXIS];
    T wr[M + 2];
    int fr[M + 2];  // bit 0: velocity BC free, bit 1: temperature BC free
#pragma unroll
    for (int j = 0; j < M + 2; j++) {
        const int r = r0 + j - 1;
        const int64_t c = (ok && r >= 0 && r < n) ? base + (int64_t)r * s : 0;
        wr[j] = wv[c];
        fr[j] = (bcv[c] == BC_FREE ? 1 : 0) | (bct[c] == BC_FREE ? 2 : 0);
    }
    const bool hasx = xs.f[0] != nullptr;  // xbar = xscale * xs, or zero
    T xr[4][M], tr[4][M];
#pragma unroll
    for (int j = 0; j < M; j++) {
        const int64_t c = (ok && r0 + j < n) ? base + (int64_t)(r0 + j) * s : 0;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            xr[v][j] = (hasx ? xs.f[v] : tb.f[v])[c];
            tr[v][j] = tb.f[v][c];
        }
    }
#pragma unroll
    for (int j = 0; j < M + 2; j++) asm volatile("" ::"v"(wr[j]), "v"(fr[j]));
    int role[M + 2];
#pragma unroll
    for (int j = 0; j < M + 2; j++) role[j] = row_role(ty[j + 1], ty[j], ty[j + 2]);
    auto coef = [&](int j) {  // row r0 + j, j in [-1, M]
        const int rl = role[j + 1];
        return row_coef<T>(rl, fr[j + 1] & 1, fr[j + 1] & 2,
                           rl == RL_IN ? wr[j + 1] * p.inv_2h : (T)0, p);
    };

    Chunk<T, M> ch;
    Row<T> rfirst, rlast;
    Coef<T> cm = coef(-1), c0 = coef(0);
#pragma unroll
    for (int j = 0; j < M; j++) {
        const Coef<T> cp = coef(j + 1);
        Row<T> r;
        r.av = cm.cv; r.bv = c0.bv; r.cv = cp.av;
        r.at = cm.ct; r.bt = c0.bt; r.ct = cp.at;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const T xv = hasx ? xscale * xr[v][j] : (T)0;
            r.d[v] = role[j + 1] == RL_IN ? xv + beta * tr[v][j] : (T)0;
        }
        if (j == 0) rfirst = r;
        else if (j == M - 1) rlast = r;
        else chunk_down<T, M>(ch, j, r);
        cm = c0;
        c0 = cp;
    }
    if (M >= 3) chunk_up<T, M>(ch);
    {
        T f[kRedFields], l[kRedFields];
        chunk_reduced<T, M>(ch, rfirst, rlast, f, l);
#pragma unroll
        for (int k = 0; k < kRedFields; k++) {
            red[L.at(gl, k, 2 * chunk)] = f[k];
            red[L.at(gl, k, 2 * chunk + 1)] = l[k];
        }
    }
    __syncthreads();
    solve_reduced_compact<T>(red, L, LPG, wave, nw);
    __syncthreads();
    if (!ok) return;
    T xf[4], xl[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        xf[v] = red[L.at(gl, 6 + v, 2 * chunk)];
        xl[v] = red[L.at(gl, 6 + v, 2 * chunk + 1)];
    }
#pragma unroll
    for (int j = 0; j < M; j++) {
        const int rl = role[j + 1];
        if (rl == RL_NONE) continue;
        const int64_t c = base + (int64_t)(r0 + j) * s;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            T x;
            if (j == 0) x = xf[v];
            else if (j == M - 1) x = xl[v];
            else {
                const T F = v < 3 ? ch.Fv[j] : ch.Ft[j];
                const T G = v < 3 ? ch.Gv[j] : ch.Gt[j];
                x = ch.E[v][j] + F * xf[v] + G * xl[v];
            }
            lam.f[v][c] = x;
        }
    }
}

// Transposed line solve along z in the contiguous layout of the forward Z
// sweep: 32 lanes per line, lane l owns rows [l*M, l*M+M) (16-byte vector
// accesses), two lines per wavefront, kZWaves wavefronts per work-group.  The
// coefficients of the rows either side of a chunk come from the neighbouring
// lanes by shuffle (c of the previous row, a of the next row).
template <class T, int M, bool VALL>
__global__ void __launch_bounds__(64 * kZWaves) adj_z_kernel(
    Geom g, SweepParams<T> p, const uint8_t* __restrict__ type, const uint8_t* __restrict__ bcv,
    const uint8_t* __restrict__ bct, Layer4<T> tin, Layer4<T> xs, T xscale, Layer4<T> tb, T beta,
    Layer4<T> lam, Layer4<T> cacc, int zacc, const int32_t* __restrict__ lines) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const int n = g.nz;
    const int used = (n + M - 1) / M;
    const int sub = lane >> 5, ll = lane & 31;
    const int b = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int lid = lines[b * (2 * nw) + 2 * wave + sub];
    const bool live = lid >= 0;
    const int64_t base = live ? (int64_t)lid * g.nz : 0;
    const int r0 = ll * M;
    const bool ok = live && ll < used;
    const int valid = ok ? min(M, n - r0) : 0;
    const bool vec = VALL && valid == M;
    const int64_t c0 = base + r0;

    // Every load of the lane goes out in one batch (a value used under a
    // condition only would be loaded there, behind its own wait): types, BC
    // kinds, the advecting velocity, xbar, tbar and the accumulator of the
    // lane's rows; the raw values are then used once, all together.
    int tyr[M], bvr[M], btr[M];
    load_bytes_raw<M, VALL>(type, c0, valid, tyr);
    load_bytes_raw<M, VALL>(bcv, c0, valid, bvr);
    load_bytes_raw<M, VALL>(bct, c0, valid, btr);
    T w[M];
    load_run_raw<T, M, VALL>(tin.f[2], c0, valid, w);
    const bool hasx = xs.f[0] != nullptr;  // xbar = xscale * xs, or zero
    T xa[4][M], ta[4][M], ca[4][M];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        load_run_raw<T, M, VALL>(hasx ? xs.f[v] : tb.f[v], c0, valid, xa[v]);
        load_run_raw<T, M, VALL>(tb.f[v], c0, valid, ta[v]);
        // the state adjoint's accumulator (the input of a Z sweep is the
        // state): (3/dt) lam is added to it on NODE_IN rows (zacc 2) or
        // starts it (zacc 1)
        load_run_raw<T, M, VALL>(zacc == 2 ? cacc.f[v] : tb.f[v], c0, valid, ca[v]);
    }
#pragma unroll
    for (int j = 0; j < M; j++) {
        pin(tyr[j]);
        pin(bvr[j]);
        pin(btr[j]);
        pin(w[j]);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            pin(xa[v][j]);
            pin(ta[v][j]);
            pin(ca[v][j]);
        }
    }

    uint8_t ty[M + 2];
#pragma unroll
    for (int j = 0; j < M; j++) ty[j + 1] = j < valid ? (uint8_t)tyr[j] : kOutside;
    {
        const int prev = __shfl((int)ty[M], lane - 1, 64);
        const int nxt = __shfl((int)ty[1], lane + 1, 64);
        ty[0] = (ll > 0 && ok) ? (uint8_t)prev : kOutside;
        ty[M + 1] = (ll < 31 && ok && r0 + M < n) ? (uint8_t)nxt : kOutside;
    }
    int role[M];
#pragma unroll
    for (int j = 0; j < M; j++) role[j] = row_role(ty[j + 1], ty[j], ty[j + 2]);
    Coef<T> k[M];
#pragma unroll
    for (int j = 0; j < M; j++)
        k[j] = row_coef<T>(role[j], bvr[j] == BC_FREE, btr[j] == BC_FREE,
                           role[j] == RL_IN ? w[j] * p.inv_2h : (T)0, p);
    // c of row r0-1 and a of row r0+M (zero outside the line: identity rows)
    T cvm = __shfl(k[M - 1].cv, lane - 1, 64), ctm = __shfl(k[M - 1].ct, lane - 1, 64);
    T avp = __shfl(k[0].av, lane + 1, 64), atp = __shfl(k[0].at, lane + 1, 64);
    if (!(ll > 0 && ok)) cvm = ctm = 0;
    if (!(ll < 31 && ok && r0 + M < n)) avp = atp = 0;

    T rhs[4][M];
#pragma unroll
    for (int v = 0; v < 4; v++)
#pragma unroll
        for (int j = 0; j < M; j++) {
            const T xv = hasx ? xscale * xa[v][j] : (T)0;
            rhs[v][j] = role[j] == RL_IN ? xv + beta * ta[v][j] : (T)0;
            if (zacc != 2) ca[v][j] = 0;
        }
    Chunk<T, M> ch;
    Row<T> rfirst, rlast;
#pragma unroll
    for (int j = 0; j < M; j++) {
        Row<T> r;
        r.av = j == 0 ? cvm : k[j - 1].cv;
        r.at = j == 0 ? ctm : k[j - 1].ct;
        r.bv = k[j].bv;
        r.bt = k[j].bt;
        r.cv = j == M - 1 ? avp : k[j + 1].av;
        r.ct = j == M - 1 ? atp : k[j + 1].at;
#pragma unroll
        for (int v = 0; v < 4; v++) r.d[v] = rhs[v][j];
        if (j == 0) rfirst = r;
        else if (j == M - 1) rlast = r;
        else chunk_down<T, M>(ch, j, r);
    }
    if (M >= 3) chunk_up<T, M>(ch);
    // lanes without rows hold identity rows; the line's reduced system stays in
    // its 32 lanes
    T xf[4], xl[4];
    {
        T f[kRedFields], l[kRedFields];
        chunk_reduced<T, M>(ch, rfirst, rlast, f, l);
        solve_chunks_in_lanes<T>(f, l, ll, used, 32, xf, xl);
    }
    if (!ok) return;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        T x[M];
#pragma unroll
        for (int j = 0; j < M; j++) {
            if (j == 0) x[j] = xf[v];
            else if (j == M - 1) x[j] = xl[v];
            else {
                const T F = v < 3 ? ch.Fv[j] : ch.Ft[j];
                const T G = v < 3 ? ch.Gv[j] : ch.Gt[j];
                x[j] = ch.E[v][j] + F * xf[v] + G * xl[v];
            }
        }
        store_run<T, M>(lam.f[v] + c0, x, vec, valid);
        if (zacc) {
#pragma unroll
            for (int j = 0; j < M; j++)
                if (role[j] == RL_IN) ca[v][j] = zacc == 2 ? ca[v][j] + p.three_dt * x[j] : p.three_dt * x[j];
            store_run<T, M>(cacc.f[v] + c0, ca[v], vec, valid);
        }
    }
}

// Node word of the pointwise kernel, constant over a backward pass (built by
// adj_mask_kernel at its start): bits 0-1 the node type, bit 2 / 3 the
// velocity / temperature BC is free, bit 4 + nb the neighbour q - e_o (nb =
// 2 o) or q + e_o (nb = 2 o + 1) lies inside the grid and is NODE_IN.  One
// 16-bit load instead of nine byte loads per node.
constexpr int kMaskFreeV = 4, kMaskFreeT = 8, kMaskNb = 4;

__global__ void __launch_bounds__(256) adj_mask_kernel(Geom g, const uint8_t* __restrict__ type,
                                                       const uint8_t* __restrict__ bcv,
                                                       const uint8_t* __restrict__ bct,
                                                       uint16_t* __restrict__ mask) {
    const int sx = (int)g.sx, sy = (int)g.sy;
    const int so[3] = {sx, sy, 1};
    const int dim[3] = {g.nx, g.ny, g.nz};
    const uint32_t N = (uint32_t)g.N;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < N; q += gridDim.x * blockDim.x) {
        const int ux = (int)(q / (uint32_t)sx), rem = (int)q - ux * sx, uy = rem / sy;
        const int pos[3] = {ux, uy, rem - uy * sy};
        uint32_t m = type[q] & 3u;
        if (bcv[q] == BC_FREE) m |= kMaskFreeV;
        if (bct[q] == BC_FREE) m |= kMaskFreeT;
        for (int nb = 0; nb < 6; nb++) {
            const int o = nb >> 1, k = (nb & 1) ? 1 : -1;
            if ((unsigned)(pos[o] + k) < (unsigned)dim[o] && type[(int)q + k * so[o]] == NODE_IN)
                m |= 1u << (kMaskNb + nb);
        }
        mask[q] = (uint16_t)m;
    }
}

// Pointwise part of a reversed sweep (see the file comment).  A fixed grid, so
// that block b always owns partial-sum slot b.  Each XCD (blocks b mod 8)
// takes one contiguous eighth of the nodes and its blocks sweep it together,
// one window of (blocks per XCD) x 256 consecutive nodes after another, so
// that the stencil neighbours a block reads (x-planes either side) are read by
// the XCD's other blocks at about the same time and stay in that XCD's L2.
// Every load of a node goes out in one batch: an address is q + an offset of
// at most two nodes per axis, replaced by q itself when it leaves the grid (so
// no address waits for a node type), and the terms that the node types rule
// out are dropped by selects; the stores come last.
//
// With a the sweep axis and b, c the other two, a NODE_IN row r reads the
// dissipation phi(r) = sum_f (D_a F_f)^2 + (D_a F_a)^2 + sum_{o != a} D_a F_o D_o F_a
// (F = U, V, W; D_o the central difference along o at r), so node q collects,
// from the NODE_IN rows r = q -+ e_o (sign +-1):
//   o = a:  F_a += 4 D_a F_a / 2h_a,  F_b += (2 D_a F_b + D_b F_a) / 2h_a  (and
//           c alike), all times t_phi lam_T(r); T += -v_T lam_a(r) / 2h_a
//   o != a: F_a += t_phi lam_T(r) D_a F_o(r) / 2h_o.
// The node-value gradient of the sweep's boundary rows (lam on the rows whose
// RHS is a node value: Dirichlet ends and shared rows) is added here too.
template <class T, int AXIS>
__global__ void __launch_bounds__(256) adj_point_kernel(Geom g, SweepParams<T> p, T inv_h2,
                                                        const uint16_t* __restrict__ nmask,
                                                        Layer4<T> tin, Layer4<T> out, Layer4<T> lam,
                                                        Layer4<T> tb, Layer4<T> xb, Layer4<T> cacc,
                                                        Layer4<T> gb, int last, T alpha,
                                                        double* __restrict__ pp) {
    constexpr int A = AXIS;
    // a grid holds fewer than 2^31 nodes (checked at creation): 32-bit indices
    const int sx = (int)g.sx, sy = (int)g.sy;
    const int so[3] = {sx, sy, 1};
    const int dim[3] = {g.nx, g.ny, g.nz};
    const T io[3] = {p.inv_2hx, p.inv_2hy, p.inv_2hz};
    const T ia = io[A];
    const T* F[3] = {tin.f[0], tin.f[1], tin.f[2]};
    const T* Tt = tin.f[3];
    double part[4] = {0, 0, 0, 0};
    const uint32_t N = (uint32_t)g.N;
    // the grid is a multiple of 8 blocks (the launch sizes it so)
    const uint32_t nbx = gridDim.x / 8, xcd = blockIdx.x % 8, local = blockIdx.x / 8;
    const uint32_t slab = ((N + 7) / 8 + blockDim.x - 1) / blockDim.x * blockDim.x;
    const uint32_t q1 = min(N, (xcd + 1) * slab);
    for (uint32_t q = xcd * slab + local * blockDim.x + threadIdx.x; q < q1; q += nbx * blockDim.x) {
        const int ux = (int)(q / (uint32_t)sx), rem = (int)q - ux * sx, uy = rem / sy;
        const int pos[3] = {ux, uy, rem - uy * sy};
        // q + k1 e_o1 + k2 e_o2 inside the grid, else q (only the axes that
        // move are checked; the offset is masked, not selected: a select
        // between two addresses becomes a select between two loads, one of
        // them the load at q already made, and the other a conditional load)
        auto inr = [&](int o, int k) { return (unsigned)(pos[o] + k) < (unsigned)dim[o]; };
        auto at = [&](int o1, int k1, int o2 = 0, int k2 = 0) -> uint32_t {
            bool in;
            if (k2 == 0) in = k1 == 0 || inr(o1, k1);
            else if (o1 == o2) in = k1 + k2 == 0 || inr(o1, k1 + k2);
            else in = (k1 == 0 || inr(o1, k1)) && inr(o2, k2);
            const int off = k1 * so[o1] + k2 * so[o2];
            return (uint32_t)((int)q + (off & -(int)in));
        };
        const int mq = nmask[q];  // node type, BC kinds, NODE_IN neighbours
        // node-value gradient and state-adjoint accumulator, read whatever the
        // node types (their updates below are conditional)
        T gbq[4], caq[4];
#pragma unroll
        for (int f = 0; f < 4; f++) {
            gbq[f] = gb.f[f][q];
            caq[f] = last ? cacc.f[f][q] : (T)0;
        }

        // own row: lam, the sweep output around q, the T and velocity stencils
        T tbq[4], l[4], xm[4], x0[4], xp[4];
#pragma unroll
        for (int f = 0; f < 4; f++) {
            tbq[f] = tb.f[f][q];
            l[f] = lam.f[f][q];
            xm[f] = out.f[f][at(A, -1)];
            x0[f] = out.f[f][q];
            xp[f] = out.f[f][at(A, 1)];
        }
        const T tgrad = (Tt[at(A, 1)] - Tt[at(A, -1)]) * p.inv_2h;
        T da[3], dA[3];  // D_a F_f and D_o F_a at q
#pragma unroll
        for (int f = 0; f < 3; f++) da[f] = (F[f][at(A, 1)] - F[f][at(A, -1)]) * ia;
#pragma unroll
        for (int o = 0; o < 3; o++)
            if (o != A) dA[o] = (F[A][at(o, 1)] - F[A][at(o, -1)]) * io[o];

        // neighbour rows r = q + k e_o: t_phi lam_T(r) and the differences
        // their stencils need
        T kr[6], lr[6], ra[6][3], rA[6][3];
#pragma unroll
        for (int nb = 0; nb < 6; nb++) {
            const int o = nb >> 1, k = (nb & 1) ? 1 : -1;
            const uint32_t r = at(o, k);
            kr[nb] = p.t_phi * lam.f[3][r];
            if (o == A) {
                lr[nb] = lam.f[A][r];
#pragma unroll
                for (int f = 0; f < 3; f++)
                    ra[nb][f] = (F[f][at(A, k + 1)] - F[f][at(A, k - 1)]) * ia;
#pragma unroll
                for (int o2 = 0; o2 < 3; o2++)
                    if (o2 != A) rA[nb][o2] = (F[A][at(A, k, o2, 1)] - F[A][at(A, k, o2, -1)]) * io[o2];
            } else {
                ra[nb][o] = (F[o][at(o, k, A, 1)] - F[o][at(o, k, A, -1)]) * ia;
            }
        }
        // every value above is used here, once, all together: a value used
        // under a condition only would be loaded there, behind its own wait
        asm volatile("" ::"v"(mq));
#pragma unroll
        for (int nb = 0; nb < 6; nb++) asm volatile("" ::"v"(kr[nb]));
#pragma unroll
        for (int f = 0; f < 4; f++)
            asm volatile("" ::"v"(tbq[f]), "v"(l[f]), "v"(xm[f]), "v"(x0[f]), "v"(xp[f]),
                         "v"(gbq[f]), "v"(caq[f]));
        asm volatile("" ::"v"(tgrad));
#pragma unroll
        for (int f = 0; f < 3; f++) asm volatile("" ::"v"(da[f]));
#pragma unroll
        for (int o = 0; o < 3; o++)
            if (o != A) asm volatile("" ::"v"(dA[o]));
#pragma unroll
        for (int nb = 0; nb < 6; nb++) {
            const int o = nb >> 1;
            if (o == A) {
                asm volatile("" ::"v"(lr[nb]));
#pragma unroll
                for (int f = 0; f < 3; f++) asm volatile("" ::"v"(ra[nb][f]));
#pragma unroll
                for (int o2 = 0; o2 < 3; o2++)
                    if (o2 != A) asm volatile("" ::"v"(rA[nb][o2]));
            } else {
                asm volatile("" ::"v"(ra[nb][o]));
            }
        }
        const int tq = mq & 3;
        const bool inq = tq == NODE_IN;
        // the rows either side: NODE_IN or not (kOutside stands for any other)
        uint8_t tn[6];
#pragma unroll
        for (int nb = 0; nb < 6; nb++) tn[nb] = (mq >> (kMaskNb + nb)) & 1 ? (uint8_t)NODE_IN : kOutside;
        const bool doxg = last && (tq == NODE_BOUND || tq == NODE_VALVE);

        T acc[4];
#pragma unroll
        for (int f = 0; f < 4; f++) acc[f] = inq ? alpha * tbq[f] : tbq[f];
        if (inq) {
            const T abv = -(l[0] * xm[0] + l[1] * xm[1] + l[2] * xm[2]);
            const T bbv = -(l[0] * x0[0] + l[1] * x0[1] + l[2] * x0[2]);
            const T cbv = -(l[0] * xp[0] + l[1] * xp[1] + l[2] * xp[2]);
            const T abt = -l[3] * xm[3], bbt = -l[3] * x0[3], cbt = -l[3] * xp[3];
            acc[A] += (-abv + cbv - abt + cbt) * p.inv_2h;
            part[1] += (double)((-abv - cbv + 2 * bbv) * inv_h2);
            part[2] += (double)((-abt - cbt + 2 * bbt) * inv_h2);
            part[0] += (double)(-l[A] * tgrad);
            T phi = 0;
#pragma unroll
            for (int f = 0; f < 3; f++) phi += f == A ? 2 * da[f] * da[f] : da[f] * da[f];
#pragma unroll
            for (int o = 0; o < 3; o++)
                if (o != A) phi += da[o] * dA[o];
            part[3] += (double)(l[3] * phi);
        }
        // transposed stencils of the rows that read node q: row r = q - e
        // contributes with +, row r = q + e with -
#pragma unroll
        for (int nb = 0; nb < 6; nb++) {
            const int o = nb >> 1;
            const T fsg = (nb & 1) ? (T)-1 : (T)1;
            const T k = kr[nb];
            const bool ok = tn[nb] == NODE_IN;
            if (o == A) {
#pragma unroll
                for (int f = 0; f < 3; f++) {
                    const T d = f == A ? k * 4 * ra[nb][f] * ia
                                       : k * (2 * ra[nb][f] + rA[nb][f]) * ia;
                    acc[f] += ok ? fsg * d : (T)0;
                }
                acc[3] += ok ? fsg * (-p.v_T * lr[nb] * p.inv_2h) : (T)0;
            } else {
                acc[A] += ok ? fsg * (k * ra[nb][o] * io[o]) : (T)0;
            }
        }

        // node-value gradient: lam on the sweep's boundary rows whose RHS is the
        // node value (both matrices separately), and at the end of a reversed
        // step tbar on BOUND/VALVE nodes
        const int rl = row_role((uint8_t)tq, tn[2 * A], tn[2 * A + 1]);
        const bool bnd = rl != RL_IN && rl != RL_NONE;
        const bool ends = rl == RL_START || rl == RL_END;
        const bool gv = bnd && (!ends || !(mq & kMaskFreeV)), gt = bnd && (!ends || !(mq & kMaskFreeT));
#pragma unroll
        for (int f = 0; f < 4; f++) {
            const bool gl = f < 3 ? gv : gt;
            if (gl || doxg) gb.f[f][q] = (gbq[f] + (gl ? l[f] : (T)0)) + (doxg ? acc[f] : (T)0);
        }

        // stores: tbar (the input adjoint of an X or Y sweep stays in lam for the
        // next solve; that of a Z sweep went to the accumulator in its solve)
#pragma unroll
        for (int f = 0; f < 4; f++) {
            if (last) xb.f[f][q] = inq ? caq[f] + acc[f] : (T)0;
            // end of the reversed step (the Picard layer started as the
            // state, whose BOUND/VALVE nodes UpdateBoundaries filled with
            // node values): xbar of the previous step, tbar cleared
            tb.f[f][q] = last ? (T)0 : acc[f];
        }
    }
    // block sums of the parameter partials, added to this block's slot
    __shared__ double red[4][256 / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        double v = part[i];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) red[i][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) v += red[threadIdx.x][w];
        pp[4 * blockIdx.x + threadIdx.x] += v;
    }
}

int adj_z_chunk(int n) {
    for (int m : {2, 4, 8})
        if (32 * m >= n) return m;
    return 0;
}

int adj_waves(int n, int lpg) {
    const int gpw = 64 / lpg;
    return (n + kAdjChunk * gpw - 1) / (kAdjChunk * gpw);
}

}  // namespace

// Lines per tile of the transposed X/Y solves: 16 (64-byte FP32 rows).  A
// work-group then holds a line's rows in ceil(n / 32) waves (3 for 96 nodes):
// with the ~150 registers the solve takes, a SIMD holds 3 waves, and
// work-groups of 6 waves (32 lines) run one per CU (their waves take two
// slots on two SIMDs), so a grid of more tiles than CUs ran in two rounds;
// 3-wave work-groups run 2-3 per CU, all tiles of a 96^3 grid at once.
template <class T>
int adj_lpg(const Geom&) {
    return 16;
}

template <class T>
bool adjoint_supported(const Geom& g) {
    const int lpg = adj_lpg<T>(g);
    for (int n : {g.nx, g.ny})
        if (adj_waves(n, lpg) > 8) return false;
    return adj_z_chunk(g.nz) != 0;
}

void build_adj_tiles(const uint8_t* t, const Geom& g, int axis, int lpg, int32_t** d_tiles,
                     int64_t* ntiles, int* ltiles) {
    std::vector<int32_t> items;
    if (axis == AX_Z) {
        // z lines x*ny + y holding a NODE_IN node, padded with -1 to whole work-groups
        for (int x = 0; x < g.nx; x++)
            for (int y = 0; y < g.ny; y++) {
                const uint8_t* row = &t[x * g.sx + y * g.sy];
                bool any = false;
                for (int z = 0; z < g.nz && !any; z++) any = row[z] == NODE_IN;
                if (any) items.push_back(x * g.ny + y);
            }
        while (items.size() % (2 * kZWaves)) items.push_back(-1);
        *ntiles = (int64_t)items.size() / (2 * kZWaves);  // work-groups
        *ltiles = 0;
        *d_tiles = nullptr;
        if (!items.empty()) {
            CMCX_HIP(hipMalloc((void**)d_tiles, sizeof(int32_t) * items.size()));
            CMCX_HIP(hipMemcpy(*d_tiles, items.data(), sizeof(int32_t) * items.size(),
                               hipMemcpyHostToDevice));
        }
        return;
    }
    // lines of a tile: z lines (X, Y axes) or y lines (Z axis)
    const int nl = axis == AX_Z ? g.ny : g.nz;
    const int lt = (nl + lpg - 1) / lpg;
    const int nouter = axis == AX_X ? g.ny : g.nx;
    const int n = axis == AX_X ? g.nx : axis == AX_Y ? g.ny : g.nz;
    for (int o = 0; o < nouter; o++)
        for (int k = 0; k < lt; k++) {
            bool any = false;
            for (int li = k * lpg; li < std::min(nl, (k + 1) * lpg) && !any; li++)
                for (int r = 0; r < n && !any; r++) {
                    int64_t c;
                    if (axis == AX_X) c = (int64_t)r * g.sx + (int64_t)o * g.sy + li;
                    else if (axis == AX_Y) c = (int64_t)o * g.sx + (int64_t)r * g.sy + li;
                    else c = (int64_t)o * g.sx + (int64_t)li * g.sy + r;
                    any = t[c] == NODE_IN;
                }
            if (any) items.push_back(o * lt + k);
        }
    *ntiles = (int64_t)items.size();
    *ltiles = lt;
    *d_tiles = nullptr;
    if (!items.empty()) {
        CMCX_HIP(hipMalloc((void**)d_tiles, sizeof(int32_t) * items.size()));
        CMCX_HIP(hipMemcpy(*d_tiles, items.data(), sizeof(int32_t) * items.size(),
                           hipMemcpyHostToDevice));
    }
}

int adj_point_blocks() { return kPointBlocks; }

void launch_adj_mask(const Geom& g, const uint8_t* type, const uint8_t* bcv, const uint8_t* bct,
                     uint16_t* mask, hipStream_t st) {
    const int64_t blocks = std::min<int64_t>((g.N + 255) / 256, 8192);
    hipLaunchKernelGGL(adj_mask_kernel, dim3((unsigned)blocks), dim3(256), 0, st, g, type, bcv, bct,
                       mask);
    CMCX_LAUNCH_CHECK();
}

namespace {

// Grid of adj_point: the blocks that are resident at once on this device at
// the kernel's occupancy (one round, no tail), at most kPointBlocks.
template <class T, int AX>
unsigned point_grid() {
    static std::atomic<int> cached[16];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) return kPointBlocks;
    int n = cached[dev].load(std::memory_order_relaxed);
    if (n == 0) {
        int per = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, adj_point_kernel<T, AX>, 256, 0) !=
                hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            per <= 0 || cus <= 0)
            n = kPointBlocks;
        else
            n = std::max(8, std::min(kPointBlocks, per * cus) / 8 * 8);
        cached[dev].store(n, std::memory_order_relaxed);
    }
    return (unsigned)n;
}

}  // namespace

template <class T>
void launch_adj_sweep(const Geom& g, int axis, const SweepParams<T>& p, const uint8_t* type,
                      const uint8_t* bcv, const uint8_t* bct, Layer4<T> tin, Layer4<T> out,
                      Layer4<T> xs, T xscale, Layer4<T> xb, Layer4<T> tb, Layer4<T> cacc,
                      Layer4<T> lam, Layer4<T> gb, T alpha, T beta, int zmode, int first_z, int last,
                      const int32_t* tiles, int64_t ntiles, int ltiles, double* pp,
                      const uint16_t* nmask, hipStream_t st) {
    const int lpg = adj_lpg<T>(g);
    if (axis == AX_Z) {
        if (ntiles > 0) {
            const int m = adj_z_chunk(g.nz);
            const size_t lds = 0;
            // whole 16-byte aligned runs: every lane's M rows inside its line
            const int vec_ok = ((size_t)g.nz * sizeof(T)) % 16 == 0 && g.nz % m == 0 ? 1 : 0;
#define LZ(MM, VA)                                                                              \
    hipLaunchKernelGGL((adj_z_kernel<T, MM, VA>), dim3((unsigned)ntiles), dim3(64 * kZWaves), lds, \
                       st, g, p, type, bcv, bct, tin, xs, xscale, tb, beta, lam, cacc,             \
                       zmode ? (first_z ? 1 : 2) : 0, tiles)
            if (vec_ok) {
                if (m == 2) LZ(2, true);
                else if (m == 4) LZ(4, true);
                else LZ(8, true);
            } else {
                if (m == 2) LZ(2, false);
                else if (m == 4) LZ(4, false);
                else LZ(8, false);
            }
#undef LZ
            CMCX_LAUNCH_CHECK();
        }
    } else if (ntiles > 0) {
        const int n = axis == AX_X ? g.nx : g.ny;
        // at most 8 (adjoint_supported); more waves than the rows need when that
        // lets the reduced solve take every line of the tile in one pass (a
        // wave solves 64 / RP lines at a time, RP = the chunk count rounded up
        // to a power of two): for 96 rows, 4 waves (one pass) instead of 3 (two)
        int waves = adj_waves(n, lpg);
        while (waves < 8) {
            int rp = 1;
            while (rp < waves * (64 / lpg)) rp <<= 1;
            if (waves * (64 / rp) >= lpg) break;
            waves++;
        }
        const int P = waves * (64 / lpg);
        const size_t lds = (size_t)lpg * (kRedFields * 2 * P + 1) * sizeof(T);
#define LA(AX, LPG)                                                                               \
    hipLaunchKernelGGL((adj_solve_kernel<T, AX, kAdjChunk, LPG, 512>), dim3((unsigned)ntiles),    \
                       dim3(64 * waves), lds, st, g, p, type, bcv, bct, tin, xs, xscale, tb, beta, lam, \
                       tiles, ltiles)
        if (axis == AX_X) LA(AX_X, 16);
        else LA(AX_Y, 16);
#undef LA
        CMCX_LAUNCH_CHECK();
    }
    const T h = axis == AX_X ? p.hx : axis == AX_Y ? p.hy : p.hz;
    const T inv_h2 = (T)1 / (h * h);
    // every launch of one kernel uses the same grid, so that slot b always belongs
    // to block b (slots past the grid stay zero)
#define LP(AX)                                                                                  \
    hipLaunchKernelGGL((adj_point_kernel<T, AX>), dim3(point_grid<T, AX>()), dim3(256), 0, st, g, p, \
                       inv_h2, nmask, tin, out, lam, tb, xb, cacc, gb, last, alpha, pp)
    if (axis == AX_X) LP(AX_X);
    else if (axis == AX_Y) LP(AX_Y);
    else LP(AX_Z);
#undef LP
    CMCX_LAUNCH_CHECK();
}

#define CMCX_ADJ_INST(T)                                                                          \
    template int adj_lpg<T>(const Geom&);                                                        \
    template bool adjoint_supported<T>(const Geom&);                                             \
    template void launch_adj_sweep<T>(const Geom&, int, const SweepParams<T>&, const uint8_t*,    \
                                      const uint8_t*, const uint8_t*, Layer4<T>, Layer4<T>,       \
                                      Layer4<T>, T, Layer4<T>, Layer4<T>, Layer4<T>, Layer4<T>,   \
                                      Layer4<T>, T, T, int, int, int, const int32_t*, int64_t,    \
                                      int, double*, const uint16_t*, hipStream_t);
CMCX_ADJ_INST(float)
CMCX_ADJ_INST(double)

}  // namespace cmcx
