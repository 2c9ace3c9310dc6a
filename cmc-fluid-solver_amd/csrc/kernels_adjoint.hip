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
//   adj_point  (pointwise)  with Mbar = -lam x^T on the tridiagonal pattern:
//              tbar <- alpha tbar (NODE_IN) + d(a, c)/dw : Mbar (advecting
//              velocity) + the transposed pressure and dissipation stencils of
//              the neighbouring rows (a gather: each node collects the terms of
//              the rows that read it, so no two threads write one node);
//              inbar = (3/dt) lam (the adjoint of the sweep input, or added to
//              the state adjoint for a Z sweep, whose input is the state);
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
// term adv = w / 2h, boundary and identity rows from the role and BC kinds.
template <class T>
__device__ __forceinline__ Coef<T> row_coef(int role, bool fv, bool ft, T adv,
                                            const SweepParams<T>& p) {
    Coef<T> k;
    if (role == RL_IN) {
        k.av = -adv - p.vis_v;
        k.cv = adv - p.vis_v;
        k.bv = p.three_dt + 2 * p.vis_v;
        k.at = -adv - p.vis_t;
        k.ct = adv - p.vis_t;
        k.bt = p.three_dt + 2 * p.vis_t;
    } else {
        bool u;
        bc_row<T>(role, role != RL_NONE && fv, k.av, k.bv, k.cv, u);
        bc_row<T>(role, role != RL_NONE && ft, k.at, k.bt, k.ct, u);
    }
    return k;
}

// Transposed line solve.  Work-group = LPG neighbouring lines x all rows, in the
// layout of the strided forward kernel; for the Z axis the lines of a tile are
// LPG consecutive y at one x (rows contiguous in memory).
template <class T, int AXIS, int M, int LPG, int MAXT>
__global__ void __launch_bounds__(MAXT) adj_solve_kernel(
    Geom g, SweepParams<T> p, const uint8_t* __restrict__ type, const uint8_t* __restrict__ bcv,
    const uint8_t* __restrict__ bct, Layer4<T> tin, Layer4<T> xb, Layer4<T> tb, T beta,
    Layer4<T> lam, Layer4<T> gb, const int32_t* __restrict__ tiles, int ltiles) {
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

    // node types of rows r0-2 .. r0+M+1, roles of rows r0-1 .. r0+M
    uint8_t ty[M + 4];
#pragma unroll
    for (int j = 0; j < M + 4; j++) {
        const int r = r0 + j - 2;
        ty[j] = (ok && r >= 0 && r < n) ? type[base + (int64_t)r * s] : kOutside;
    }
    int role[M + 2];
#pragma unroll
    for (int j = 0; j < M + 2; j++) role[j] = row_role(ty[j + 1], ty[j], ty[j + 2]);
    const T* wv = tin.f[AXIS];
    auto coef = [&](int j) {  // row r0 + j, j in [-1, M]
        const int rl = role[j + 1];
        const int64_t c = base + (int64_t)(r0 + j) * s;
        T adv = 0;
        bool fv = false, ft = false;
        if (rl == RL_IN) adv = wv[c] * p.inv_2h;
        else if (rl != RL_NONE) {
            fv = bcv[c] == BC_FREE;
            ft = bct[c] == BC_FREE;
        }
        return row_coef<T>(rl, fv, ft, adv, p);
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
        const int64_t c = base + (int64_t)(r0 + j) * s;
        if (role[j + 1] == RL_IN) {
#pragma unroll
            for (int v = 0; v < 4; v++) r.d[v] = xb.f[v][c] + beta * tb.f[v][c];
        } else {
#pragma unroll
            for (int v = 0; v < 4; v++) r.d[v] = 0;
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
        bool uv = rl != RL_IN, ut = rl != RL_IN;
        if (rl == RL_START || rl == RL_END) {
            uv = bcv[c] != BC_FREE;
            ut = bct[c] != BC_FREE;
        }
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
            if ((v < 3 && uv) || (v == 3 && ut)) gb.f[v][c] += x;
        }
    }
}

// Transposed line solve along z in the contiguous layout of the forward Z
// sweep: 32 lanes per line, lane l owns rows [l*M, l*M+M) (16-byte vector
// accesses), two lines per wavefront, kZWaves wavefronts per work-group.  The
// coefficients of the rows either side of a chunk come from the neighbouring
// lanes by shuffle (c of the previous row, a of the next row).
template <class T, int M>
__global__ void __launch_bounds__(64 * kZWaves) adj_z_kernel(
    Geom g, SweepParams<T> p, const uint8_t* __restrict__ type, const uint8_t* __restrict__ bcv,
    const uint8_t* __restrict__ bct, Layer4<T> tin, Layer4<T> xb, Layer4<T> tb, T beta,
    Layer4<T> lam, Layer4<T> gb, const int32_t* __restrict__ lines, int vec_ok) {
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
    const bool vec = vec_ok && valid == M;
    const int64_t c0 = base + r0;

    uint8_t ty[M + 2];
#pragma unroll
    for (int j = 0; j < M; j++) ty[j + 1] = j < valid ? type[c0 + j] : kOutside;
    {
        const int prev = __shfl((int)ty[M], lane - 1, 64);
        const int nxt = __shfl((int)ty[1], lane + 1, 64);
        ty[0] = (ll > 0 && ok) ? (uint8_t)prev : kOutside;
        ty[M + 1] = (ll < 31 && ok && r0 + M < n) ? (uint8_t)nxt : kOutside;
    }
    int role[M];
#pragma unroll
    for (int j = 0; j < M; j++) role[j] = row_role(ty[j + 1], ty[j], ty[j + 2]);
    T w[M];
    load_run<T, M>(tin.f[2] + c0, w, vec, valid);
    Coef<T> k[M];
#pragma unroll
    for (int j = 0; j < M; j++) {
        bool fv = false, ft = false;
        if (role[j] != RL_IN && role[j] != RL_NONE) {
            fv = bcv[c0 + j] == BC_FREE;
            ft = bct[c0 + j] == BC_FREE;
        }
        k[j] = row_coef<T>(role[j], fv, ft, role[j] == RL_IN ? w[j] * p.inv_2h : (T)0, p);
    }
    // c of row r0-1 and a of row r0+M (zero outside the line: identity rows)
    T cvm = __shfl(k[M - 1].cv, lane - 1, 64), ctm = __shfl(k[M - 1].ct, lane - 1, 64);
    T avp = __shfl(k[0].av, lane + 1, 64), atp = __shfl(k[0].at, lane + 1, 64);
    if (!(ll > 0 && ok)) cvm = ctm = 0;
    if (!(ll < 31 && ok && r0 + M < n)) avp = atp = 0;

    T rhs[4][M];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        T xa[M], ta[M];
        load_run<T, M>(xb.f[v] + c0, xa, vec, valid);
        load_run<T, M>(tb.f[v] + c0, ta, vec, valid);
#pragma unroll
        for (int j = 0; j < M; j++) rhs[v][j] = role[j] == RL_IN ? xa[j] + beta * ta[j] : (T)0;
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
#pragma unroll
        for (int j = 0; j < M; j++) {
            const int rl = role[j];
            if (rl == RL_IN || rl == RL_NONE || j >= valid) continue;
            bool use = true;
            if (rl == RL_START || rl == RL_END) use = (v < 3 ? bcv[c0 + j] : bct[c0 + j]) != BC_FREE;
            if (use) gb.f[v][c0 + j] += x[j];
        }
    }
}

// Pointwise part of a reversed sweep (see the file comment).  Grid-stride over
// all nodes with a fixed grid, so that block b always owns partial-sum slot b.
template <class T, int AXIS>
__global__ void __launch_bounds__(256) adj_point_kernel(Geom g, SweepParams<T> p, T inv_h2,
                                                        const uint8_t* __restrict__ type,
                                                        Layer4<T> tin, Layer4<T> out, Layer4<T> lam,
                                                        Layer4<T> tb, Layer4<T> xb, Layer4<T> cacc,
                                                        Layer4<T> gb, int zmode, int last, T alpha,
                                                        double* __restrict__ pp) {
    // a grid holds fewer than 2^31 nodes (checked at creation): 32-bit indices
    const uint32_t sx = (uint32_t)g.sx, sy = (uint32_t)g.sy;
    const uint32_t s = AXIS == AX_X ? sx : AXIS == AX_Y ? sy : 1u;
    const T ix = p.inv_2hx, iy = p.inv_2hy, iz = p.inv_2hz;
    const T* U = tin.f[0];
    const T* V = tin.f[1];
    const T* W = tin.f[2];
    const T* Tt = tin.f[3];
    double part[4] = {0, 0, 0, 0};
    const uint32_t N = (uint32_t)g.N;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < N; q += gridDim.x * blockDim.x) {
        const uint32_t ux = q / sx, rem = q - ux * sx, uy = rem / sy;
        const int x = (int)ux, y = (int)uy, z = (int)(rem - uy * sy);
        const uint8_t tq = type[q];
        const bool inq = tq == NODE_IN;
        T acc[4], acc_state[4] = {0, 0, 0, 0};
#pragma unroll
        for (int f = 0; f < 4; f++) acc[f] = inq ? alpha * tb.f[f][q] : tb.f[f][q];
        if (inq) {
            T l[4], xm[4], x0[4], xp[4];
#pragma unroll
            for (int f = 0; f < 4; f++) {
                l[f] = lam.f[f][q];
                xm[f] = out.f[f][q - s];
                x0[f] = out.f[f][q];
                xp[f] = out.f[f][q + s];
            }
            const T abv = -(l[0] * xm[0] + l[1] * xm[1] + l[2] * xm[2]);
            const T bbv = -(l[0] * x0[0] + l[1] * x0[1] + l[2] * x0[2]);
            const T cbv = -(l[0] * xp[0] + l[1] * xp[1] + l[2] * xp[2]);
            const T abt = -l[3] * xm[3], bbt = -l[3] * x0[3], cbt = -l[3] * xp[3];
            acc[AXIS] += (-abv + cbv - abt + cbt) * p.inv_2h;
            part[1] += (double)((-abv - cbv + 2 * bbv) * inv_h2);
            part[2] += (double)((-abt - cbt + 2 * bbt) * inv_h2);
            part[0] += (double)(-l[AXIS] * ((Tt[q + s] - Tt[q - s]) * p.inv_2h));
            T phi;
            if (AXIS == AX_X) {
                T u_x = (U[q + sx] - U[q - sx]) * ix, v_x = (V[q + sx] - V[q - sx]) * ix;
                T w_x = (W[q + sx] - W[q - sx]) * ix, u_y = (U[q + sy] - U[q - sy]) * iy;
                T u_z = (U[q + 1] - U[q - 1]) * iz;
                phi = 2 * u_x * u_x + v_x * v_x + w_x * w_x + v_x * u_y + w_x * u_z;
            } else if (AXIS == AX_Y) {
                T u_y = (U[q + sy] - U[q - sy]) * iy, v_y = (V[q + sy] - V[q - sy]) * iy;
                T w_y = (W[q + sy] - W[q - sy]) * iy, v_x = (V[q + sx] - V[q - sx]) * ix;
                T v_z = (V[q + 1] - V[q - 1]) * iz;
                phi = u_y * u_y + 2 * v_y * v_y + w_y * w_y + u_y * v_x + w_y * v_z;
            } else {
                T u_z = (U[q + 1] - U[q - 1]) * iz, v_z = (V[q + 1] - V[q - 1]) * iz;
                T w_z = (W[q + 1] - W[q - 1]) * iz, w_x = (W[q + sx] - W[q - sx]) * ix;
                T w_y = (W[q + sy] - W[q - sy]) * iy;
                phi = u_z * u_z + v_z * v_z + 2 * w_z * w_z + u_z * w_x + v_z * w_y;
            }
            part[3] += (double)(l[3] * phi);
#pragma unroll
            for (int f = 0; f < 4; f++) {
                const T ib = p.three_dt * l[f];
                if (last) {
                    // the state adjoint is completed below with tbar
                    acc_state[f] = cacc.f[f][q] + ib;
                } else if (zmode) {
                    cacc.f[f][q] += ib;
                    xb.f[f][q] = 0;
                } else {
                    xb.f[f][q] = ib;
                }
            }
        } else {
#pragma unroll
            for (int f = 0; f < 4; f++) xb.f[f][q] = 0;
        }
        // transposed stencils of the rows that read node q: row c = q - e
        // contributes with +, row c = q + e with -
#pragma unroll
        for (int nb = 0; nb < 6; nb++) {
            const int e = nb >> 1;
            const int sg = (nb & 1) ? 1 : -1;
            const int pe = e == 0 ? x : e == 1 ? y : z;
            const int de = e == 0 ? g.nx : e == 1 ? g.ny : g.nz;
            const uint32_t se = e == 0 ? sx : e == 1 ? sy : 1u;
            const int pc = pe + sg;
            if (pc < 0 || pc >= de) continue;
            const uint32_t c = sg < 0 ? q - se : q + se;
            // the stencils below read around c: only a NODE_IN c, never on a
            // grid face, keeps them inside the grid
            if (type[c] != NODE_IN) continue;
            const T fsg = sg < 0 ? (T)1 : (T)-1;
            const T k = p.t_phi * lam.f[3][c];
            if (AXIS == AX_X) {
                if (e == 0) {
                    const T u_x = (U[c + sx] - U[c - sx]) * ix, v_x = (V[c + sx] - V[c - sx]) * ix;
                    const T w_x = (W[c + sx] - W[c - sx]) * ix, u_y = (U[c + sy] - U[c - sy]) * iy;
                    const T u_z = (U[c + 1] - U[c - 1]) * iz;
                    acc[0] += fsg * (k * 4 * u_x * ix);
                    acc[1] += fsg * (k * (2 * v_x + u_y) * ix);
                    acc[2] += fsg * (k * (2 * w_x + u_z) * ix);
                    acc[3] += fsg * (-p.v_T * lam.f[0][c] * p.inv_2h);
                } else if (e == 1) {
                    acc[0] += fsg * (k * ((V[c + sx] - V[c - sx]) * ix) * iy);
                } else {
                    acc[0] += fsg * (k * ((W[c + sx] - W[c - sx]) * ix) * iz);
                }
            } else if (AXIS == AX_Y) {
                if (e == 1) {
                    const T u_y = (U[c + sy] - U[c - sy]) * iy, v_y = (V[c + sy] - V[c - sy]) * iy;
                    const T w_y = (W[c + sy] - W[c - sy]) * iy, v_x = (V[c + sx] - V[c - sx]) * ix;
                    const T v_z = (V[c + 1] - V[c - 1]) * iz;
                    acc[0] += fsg * (k * (2 * u_y + v_x) * iy);
                    acc[1] += fsg * (k * 4 * v_y * iy);
                    acc[2] += fsg * (k * (2 * w_y + v_z) * iy);
                    acc[3] += fsg * (-p.v_T * lam.f[1][c] * p.inv_2h);
                } else if (e == 0) {
                    acc[1] += fsg * (k * ((U[c + sy] - U[c - sy]) * iy) * ix);
                } else {
                    acc[1] += fsg * (k * ((W[c + sy] - W[c - sy]) * iy) * iz);
                }
            } else {
                if (e == 2) {
                    const T u_z = (U[c + 1] - U[c - 1]) * iz, v_z = (V[c + 1] - V[c - 1]) * iz;
                    const T w_z = (W[c + 1] - W[c - 1]) * iz, w_x = (W[c + sx] - W[c - sx]) * ix;
                    const T w_y = (W[c + sy] - W[c - sy]) * iy;
                    acc[0] += fsg * (k * (2 * u_z + w_x) * iz);
                    acc[1] += fsg * (k * (2 * v_z + w_y) * iz);
                    acc[2] += fsg * (k * 4 * w_z * iz);
                    acc[3] += fsg * (-p.v_T * lam.f[2][c] * p.inv_2h);
                } else if (e == 0) {
                    acc[2] += fsg * (k * ((U[c + 1] - U[c - 1]) * iz) * ix);
                } else {
                    acc[2] += fsg * (k * ((V[c + 1] - V[c - 1]) * iz) * iy);
                }
            }
        }
        if (last) {
            // end of the reversed step (the Picard layer started as the state,
            // whose BOUND/VALVE nodes UpdateBoundaries filled with node values):
            // xbar of the previous step, node-value gradients, cleared accumulators
#pragma unroll
            for (int f = 0; f < 4; f++) {
                if (inq) {
                    xb.f[f][q] = acc_state[f] + acc[f];
                    cacc.f[f][q] = 0;
                } else if (tq == NODE_BOUND || tq == NODE_VALVE) {
                    gb.f[f][q] += acc[f];
                }
                tb.f[f][q] = 0;
            }
        } else {
#pragma unroll
            for (int f = 0; f < 4; f++) tb.f[f][q] = acc[f];
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

template <class T>
int adj_lpg() {
    return 128 / (int)sizeof(T);
}

template <class T>
bool adjoint_supported(const Geom& g) {
    const int lpg = adj_lpg<T>();
    for (int n : {g.nx, g.ny})
        if (adj_waves(n, lpg) * 64 > 32 * lpg) return false;
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

template <class T>
void launch_adj_sweep(const Geom& g, int axis, const SweepParams<T>& p, const uint8_t* type,
                      const uint8_t* bcv, const uint8_t* bct, Layer4<T> tin, Layer4<T> out,
                      Layer4<T> xb, Layer4<T> tb, Layer4<T> cacc, Layer4<T> lam, Layer4<T> gb,
                      T alpha, T beta, int zmode, int last, const int32_t* tiles, int64_t ntiles,
                      int ltiles, double* pp, hipStream_t st) {
    constexpr int LPG = 128 / (int)sizeof(T);
    if (axis == AX_Z) {
        if (ntiles > 0) {
            const int m = adj_z_chunk(g.nz);
            const size_t lds = 0;
            const int vec_ok = ((size_t)g.nz * sizeof(T)) % 16 == 0 ? 1 : 0;
#define LZ(MM)                                                                                  \
    hipLaunchKernelGGL((adj_z_kernel<T, MM>), dim3((unsigned)ntiles), dim3(64 * kZWaves), lds, st, \
                       g, p, type, bcv, bct, tin, xb, tb, beta, lam, gb, tiles, vec_ok)
            if (m == 2) LZ(2);
            else if (m == 4) LZ(4);
            else LZ(8);
#undef LZ
            CMCX_LAUNCH_CHECK();
        }
    } else if (ntiles > 0) {
        const int n = axis == AX_X ? g.nx : g.ny;
        const int waves = adj_waves(n, LPG);
        const int P = waves * (64 / LPG);
        const size_t lds = (size_t)LPG * (kRedFields * 2 * P + 1) * sizeof(T);
#define LA(AX)                                                                                   \
    hipLaunchKernelGGL((adj_solve_kernel<T, AX, kAdjChunk, LPG, 32 * LPG>), dim3((unsigned)ntiles), \
                       dim3(64 * waves), lds, st, g, p, type, bcv, bct, tin, xb, tb, beta, lam, gb, \
                       tiles, ltiles)
        if (axis == AX_X) LA(AX_X);
        else LA(AX_Y);
#undef LA
        CMCX_LAUNCH_CHECK();
    }
    const T h = axis == AX_X ? p.hx : axis == AX_Y ? p.hy : p.hz;
    const T inv_h2 = (T)1 / (h * h);
    // every launch uses the full fixed grid, so that slot b always belongs to block b
#define LP(AX)                                                                                  \
    hipLaunchKernelGGL((adj_point_kernel<T, AX>), dim3(kPointBlocks), dim3(256), 0, st, g, p,  \
                       inv_h2, type, tin, out, lam, tb, xb, cacc, gb, zmode, last, alpha, pp)
    if (axis == AX_X) LP(AX_X);
    else if (axis == AX_Y) LP(AX_Y);
    else LP(AX_Z);
#undef LP
    CMCX_LAUNCH_CHECK();
}

#define CMCX_ADJ_INST(T)                                                                          \
    template int adj_lpg<T>();                                                                   \
    template bool adjoint_supported<T>(const Geom&);                                             \
    template void launch_adj_sweep<T>(const Geom&, int, const SweepParams<T>&, const uint8_t*,    \
                                      const uint8_t*, const uint8_t*, Layer4<T>, Layer4<T>,       \
                                      Layer4<T>, Layer4<T>, Layer4<T>, Layer4<T>, Layer4<T>, T, T, \
                                      int, int, const int32_t*, int64_t, int, double*,             \
                                      hipStream_t);
CMCX_ADJ_INST(float)
CMCX_ADJ_INST(double)

}  // namespace cmcx
