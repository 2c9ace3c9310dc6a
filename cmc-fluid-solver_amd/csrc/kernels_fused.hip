// Fused line sweeps: all four fields of a sweep in one launch, every grid line
// solved as one tridiagonal system by a chunked partition method.
//
// Line-wide system.  Row r of a line (node c) is
//   NODE_IN node            the interior row of AdiSolver3D.cpp:755-801
//                           (a = -w/2h - nu/h^2, b = 3/dt + 2 nu/h^2, c = w/2h - nu/h^2)
//   non-IN, IN node after   ApplyBC0 row (AdiSolver3D.cpp:804-827): a segment start
//   non-IN, IN node before  ApplyBC1 row (:829-852): a segment end
//   non-IN, IN on both sides the shared end/start of two segments; both reference
//                           systems hold x = node value there (no-slip), which
//                           decouples the two sides (a free shared node is
//                           routed to the per-segment kernel by the host)
//   otherwise               identity row x = 0, never written
// On a well-formed mask the line-wide matrix is block diagonal with exactly the
// reference's segments (Grid3D.cpp:62-118) as blocks, so solving it gives every
// segment's Thomas solution up to rounding.  Outputs follow UpdateSegment
// (every segment node written) and the Picard merge on NODE_IN nodes.
//
// Chunked partition method (the Thomas/PCR hybrid of Kim, Wu, Chang and Hwu,
// "A scalable tridiagonal solver for GPUs", 2011): the rows of a line are cut
// into chunks of M consecutive rows held by one lane.  A downward then upward
// elimination inside the chunk expresses every inner row through the chunk's
// first and last unknowns, x_j = E_j + F_j x_first + G_j x_last; the first and
// last rows of all chunks form a tridiagonal "reduced" system of 2 rows per
// chunk (at most 64 rows per line), staged through LDS and solved by parallel
// cyclic reduction with one reduced row per lane and cross-lane shuffles; each
// lane then reconstructs its inner rows.  The U, V and W systems share one
// matrix (same bc_vel and viscosity); T has its own.
//
// Layouts:
//   strided lines (X and Y sweeps): a work-group owns LPG consecutive z lines
//     (LPG * sizeof(T) = 128 bytes: every row access of a lane group is one
//     cache line) and all rows; lane group g of wave w owns chunk w*(64/LPG)+g.
//   contiguous lines (Z sweep): 32 lanes per line, two lines per wavefront,
//     lane l owns rows [l*M, l*M+M) (vector loads).
// A launch covers a batch of independent grids stored back to back in the flat
// field arrays: each work-group reads its grid's geometry and coefficients from
// a descriptor table (one grid is a batch of one).
// Work-groups are mapped to tiles XCD by XCD (block b runs on XCD b % 8), so
// neighbouring tiles -- which share the cross-line stencil reads -- run on the
// same XCD at about the same time and share its L2.
#include "cmcx_internal.h"

#include <algorithm>
#include <stdexcept>

namespace cmcx {

namespace {

enum { RL_NONE = 0, RL_IN = 1, RL_START = 2, RL_END = 3, RL_SHARED = 4 };
constexpr int kRedFields = 10;  // A_v B_v C_v A_t B_t C_t D_u D_v D_w D_t
constexpr uint8_t kOutside = 255;

__device__ __forceinline__ int row_role(uint8_t t, uint8_t tl, uint8_t tr) {
    if (t == NODE_IN) return RL_IN;
    bool L = tl == NODE_IN, R = tr == NODE_IN;
    if (L && R) return RL_SHARED;
    if (R) return RL_START;
    if (L) return RL_END;
    return RL_NONE;
}

// Boundary or identity row of one matrix (ApplyBC0/ApplyBC1); uses_val: the
// RHS is the node value.
template <class T>
__device__ __forceinline__ void bc_row(int role, bool free_bc, T& a, T& b, T& c, bool& uses_val) {
    uses_val = false;
    if (role == RL_START && free_bc) { a = 0; b = 2; c = -1; return; }
    if (role == RL_END && free_bc) { a = -1; b = 2; c = 0; return; }
    a = 0; b = 1; c = 0;
    uses_val = role != RL_NONE;
}

inline __device__ unsigned xcd_remap(unsigned b, unsigned nb) {
    const unsigned X = 8;
    unsigned q = nb / X, rem = nb % X;
    unsigned xcd = b % X, local = b / X;
    return xcd < rem ? xcd * (q + 1) + local : rem * (q + 1) + (xcd - rem) * q + local;
}

template <class T>
struct Row {
    T av, bv, cv, at, bt, ct;
    T d[4];
};

// Interior row at node c (stride s along the line): both matrices and the four
// right-hand sides, with the pressure term on the field along the sweep and
// the dissipation function Phi_d on T (TimeLayer3D.h:338-343, 554-588).
template <class T, int AXIS>
__device__ __forceinline__ void interior_row(const Geom& g, const SweepParams<T>& p, int64_t c,
                                             int64_t s, const Layer4<T>& in, const Layer4<T>& tin,
                                             Row<T>& r) {
    const T* tu = tin.f[0];
    const T* tv = tin.f[1];
    const T* tw = tin.f[2];
    const T* tt = tin.f[3];
    const T w = tin.f[AXIS][c];
    const T adv = w * p.inv_2h;
    r.av = -adv - p.vis_v;
    r.cv = adv - p.vis_v;
    r.bv = p.three_dt + 2 * p.vis_v;
    r.at = -adv - p.vis_t;
    r.ct = adv - p.vis_t;
    r.bt = p.three_dt + 2 * p.vis_t;
#pragma unroll
    for (int v = 0; v < 4; v++) r.d[v] = in.f[v][c] * p.three_dt;
    const T grad = (tt[c + s] - tt[c - s]) * p.inv_2h;
    r.d[AXIS] = r.d[AXIS] - p.v_T * grad;
    T phi;
    if (AXIS == AX_X) {
        T u_x = (tu[c + s] - tu[c - s]) * p.inv_2hx;
        T v_x = (tv[c + s] - tv[c - s]) * p.inv_2hx;
        T w_x = (tw[c + s] - tw[c - s]) * p.inv_2hx;
        T u_y = (tu[c + g.sy] - tu[c - g.sy]) * p.inv_2hy;
        T u_z = (tu[c + 1] - tu[c - 1]) * p.inv_2hz;
        phi = 2 * u_x * u_x + v_x * v_x + w_x * w_x + v_x * u_y + w_x * u_z;
    } else if (AXIS == AX_Y) {
        T u_y = (tu[c + s] - tu[c - s]) * p.inv_2hy;
        T v_y = (tv[c + s] - tv[c - s]) * p.inv_2hy;
        T w_y = (tw[c + s] - tw[c - s]) * p.inv_2hy;
        T v_x = (tv[c + g.sx] - tv[c - g.sx]) * p.inv_2hx;
        T v_z = (tv[c + 1] - tv[c - 1]) * p.inv_2hz;
        phi = u_y * u_y + 2 * v_y * v_y + w_y * w_y + u_y * v_x + w_y * v_z;
    } else {
        T u_z = (tu[c + s] - tu[c - s]) * p.inv_2hz;
        T v_z = (tv[c + s] - tv[c - s]) * p.inv_2hz;
        T w_z = (tw[c + s] - tw[c - s]) * p.inv_2hz;
        T w_x = (tw[c + g.sx] - tw[c - g.sx]) * p.inv_2hx;
        T w_y = (tw[c + g.sy] - tw[c - g.sy]) * p.inv_2hy;
        phi = u_z * u_z + v_z * v_z + 2 * w_z * w_z + u_z * w_x + v_z * w_y;
    }
    r.d[3] = r.d[3] + p.t_phi * phi;
}

template <class T>
__device__ __forceinline__ void boundary_row(int role, int64_t c, const uint8_t* bcv,
                                             const uint8_t* bct, const T* const* gv, Row<T>& r) {
    bool uv, ut;
    bc_row<T>(role, role != RL_NONE && bcv[c] == BC_FREE, r.av, r.bv, r.cv, uv);
    bc_row<T>(role, role != RL_NONE && bct[c] == BC_FREE, r.at, r.bt, r.ct, ut);
#pragma unroll
    for (int v = 0; v < 3; v++) r.d[v] = uv ? gv[v][c] : (T)0;
    r.d[3] = ut ? gv[3][c] : (T)0;
}

// Chunk state of one lane: inner rows 1..M-2 as x_j = E_j + F_j x_0 + G_j x_{M-1}
// (F, G per matrix, E per field).
template <class T, int M>
struct Chunk {
    T Fv[M], Gv[M], Ft[M], Gt[M];
    T E[4][M];
};

// Downward elimination, row j in 1..M-2: x_j = p_j + q_j x_0 - r_j x_{j+1},
// (p, q, r) kept in (E, F, G).
template <class T, int M>
__device__ __forceinline__ void chunk_down(Chunk<T, M>& ch, int j, const Row<T>& r) {
    if (j == 1) {
        T iv = 1 / r.bv, it = 1 / r.bt;
        ch.Fv[j] = -r.av * iv; ch.Gv[j] = r.cv * iv;
        ch.Ft[j] = -r.at * it; ch.Gt[j] = r.ct * it;
#pragma unroll
        for (int v = 0; v < 3; v++) ch.E[v][j] = r.d[v] * iv;
        ch.E[3][j] = r.d[3] * it;
    } else {
        T iv = 1 / (r.bv - r.av * ch.Gv[j - 1]);
        T it = 1 / (r.bt - r.at * ch.Gt[j - 1]);
#pragma unroll
        for (int v = 0; v < 3; v++) ch.E[v][j] = (r.d[v] - r.av * ch.E[v][j - 1]) * iv;
        ch.E[3][j] = (r.d[3] - r.at * ch.E[3][j - 1]) * it;
        ch.Fv[j] = -r.av * ch.Fv[j - 1] * iv;
        ch.Ft[j] = -r.at * ch.Ft[j - 1] * it;
        ch.Gv[j] = r.cv * iv;
        ch.Gt[j] = r.ct * it;
    }
}

// Upward elimination: (p, q, r) -> (E, F, G).
template <class T, int M>
__device__ __forceinline__ void chunk_up(Chunk<T, M>& ch) {
    ch.Gv[M - 2] = -ch.Gv[M - 2];
    ch.Gt[M - 2] = -ch.Gt[M - 2];
#pragma unroll
    for (int j = M - 3; j >= 1; j--) {
        T rv = ch.Gv[j], rt = ch.Gt[j];
#pragma unroll
        for (int v = 0; v < 3; v++) ch.E[v][j] = ch.E[v][j] - rv * ch.E[v][j + 1];
        ch.E[3][j] = ch.E[3][j] - rt * ch.E[3][j + 1];
        ch.Fv[j] = ch.Fv[j] - rv * ch.Fv[j + 1];
        ch.Ft[j] = ch.Ft[j] - rt * ch.Ft[j + 1];
        ch.Gv[j] = -rv * ch.Gv[j + 1];
        ch.Gt[j] = -rt * ch.Gt[j + 1];
    }
}

// Reduced rows of a chunk: f couples (x_{-1}, x_0, x_{M-1}), l couples
// (x_0, x_{M-1}, x_M).  Fields A_v B_v C_v A_t B_t C_t D_u D_v D_w D_t.
template <class T, int M>
__device__ __forceinline__ void chunk_reduced(const Chunk<T, M>& ch, const Row<T>& r0,
                                              const Row<T>& rl, T* f, T* l) {
    if (M >= 3) {
        f[0] = r0.av; f[1] = r0.bv + r0.cv * ch.Fv[1]; f[2] = r0.cv * ch.Gv[1];
        f[3] = r0.at; f[4] = r0.bt + r0.ct * ch.Ft[1]; f[5] = r0.ct * ch.Gt[1];
#pragma unroll
        for (int v = 0; v < 3; v++) f[6 + v] = r0.d[v] - r0.cv * ch.E[v][1];
        f[9] = r0.d[3] - r0.ct * ch.E[3][1];
        l[0] = rl.av * ch.Fv[M - 2]; l[1] = rl.bv + rl.av * ch.Gv[M - 2]; l[2] = rl.cv;
        l[3] = rl.at * ch.Ft[M - 2]; l[4] = rl.bt + rl.at * ch.Gt[M - 2]; l[5] = rl.ct;
#pragma unroll
        for (int v = 0; v < 3; v++) l[6 + v] = rl.d[v] - rl.av * ch.E[v][M - 2];
        l[9] = rl.d[3] - rl.at * ch.E[3][M - 2];
    } else {
        f[0] = r0.av; f[1] = r0.bv; f[2] = r0.cv; f[3] = r0.at; f[4] = r0.bt; f[5] = r0.ct;
        l[0] = rl.av; l[1] = rl.bv; l[2] = rl.cv; l[3] = rl.at; l[4] = rl.bt; l[5] = rl.ct;
#pragma unroll
        for (int v = 0; v < 4; v++) { f[6 + v] = r0.d[v]; l[6 + v] = rl.d[v]; }
    }
}

// Loads the chunk's rows, eliminates them and returns the two reduced rows.
// Row r of the line lives at base + r * s; rows outside [0, n) or !ok are
// identity rows.  role[] receives the row roles.
template <class T, int AXIS, int M>
__device__ __forceinline__ void chunk_phase(const Geom& g, const SweepParams<T>& p,
                                            const uint8_t* __restrict__ type,
                                            const uint8_t* __restrict__ bcv,
                                            const uint8_t* __restrict__ bct, const T* const* gv,
                                            const Layer4<T>& in, const Layer4<T>& tin,
                                            int64_t base, int64_t s, int r0, int n, bool ok,
                                            int* role, Chunk<T, M>& ch, T* f, T* l) {
    uint8_t ty[M + 2];
#pragma unroll
    for (int j = 0; j < M + 2; j++) {
        int r = r0 + j - 1;
        ty[j] = (ok && r >= 0 && r < n) ? type[base + (int64_t)r * s] : kOutside;
    }
#pragma unroll
    for (int j = 0; j < M; j++) role[j] = row_role(ty[j + 1], ty[j], ty[j + 2]);
    Row<T> rfirst, rlast;
#pragma unroll
    for (int j = 0; j < M; j++) {
        const int64_t c = base + (int64_t)(r0 + j) * s;
        Row<T> rw;
        if (role[j] == RL_IN) interior_row<T, AXIS>(g, p, c, s, in, tin, rw);
        else boundary_row<T>(role[j], c, bcv, bct, gv, rw);
        if (j == 0) rfirst = rw;
        else if (j == M - 1) rlast = rw;
        else chunk_down<T, M>(ch, j, rw);
    }
    if (M >= 3) chunk_up<T, M>(ch);
    chunk_reduced<T, M>(ch, rfirst, rlast, f, l);
}

// Writes the chunk's rows given x_first / x_last of the four fields.
template <class T, int M>
__device__ __forceinline__ void write_phase(const Chunk<T, M>& ch, const int* role, int64_t base,
                                            int64_t s, int r0, const T* xf, const T* xl,
                                            const Layer4<T>& tin, const Layer4<T>& out,
                                            const Layer4<T>& tout, int merges) {
#pragma unroll
    for (int j = 0; j < M; j++) {
        if (role[j] == RL_NONE) continue;
        const int64_t c = base + (int64_t)(r0 + j) * s;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            T x;
            if (j == 0) x = xf[v];
            else if (j == M - 1) x = xl[v];
            else {
                T F = v < 3 ? ch.Fv[j] : ch.Ft[j];
                T G = v < 3 ? ch.Gv[j] : ch.Gt[j];
                x = ch.E[v][j] + F * xf[v] + G * xl[v];
            }
            out.f[v][c] = x;
            if (role[j] == RL_IN) {
                T m = (tin.f[v][c] + x) * (T)0.5;
                if (merges > 1) m = (m + x) * (T)0.5;
                tout.f[v][c] = m;
            }
        }
    }
}

// Parallel cyclic reduction over one reduced system per group of RP lanes:
// the lane holds row i = lane % RP (rows >= R are identity, and every access
// stays inside the lane's group).  On return D holds x.
template <class T>
__device__ __forceinline__ void pcr_solve(T* A, T* B, T* C, T* D, int i, int R, int RP) {
    // A[q], B[q], C[q]: matrix q (0 vel, 1 T); D[0..2] use matrix 0, D[3] matrix 1
    if (i >= R) {
        A[0] = A[1] = C[0] = C[1] = 0;
        B[0] = B[1] = 1;
        D[0] = D[1] = D[2] = D[3] = 0;
    }
    if (i == 0) A[0] = A[1] = 0;
    if (i == R - 1) C[0] = C[1] = 0;
    const int lane = threadIdx.x & 63;
    for (int st = 1; st < RP; st <<= 1) {
        const bool hm = i - st >= 0, hp = i + st < RP;
        const int lm = hm ? lane - st : lane, lp = hp ? lane + st : lane;
        T Am[2], Bm[2], Cm[2], Ap[2], Bp[2], Cp[2], Dm[4], Dp[4];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            Am[q] = __shfl(A[q], lm, 64); Bm[q] = __shfl(B[q], lm, 64); Cm[q] = __shfl(C[q], lm, 64);
            Ap[q] = __shfl(A[q], lp, 64); Bp[q] = __shfl(B[q], lp, 64); Cp[q] = __shfl(C[q], lp, 64);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) { Dm[v] = __shfl(D[v], lm, 64); Dp[v] = __shfl(D[v], lp, 64); }
        if (!hm) {
#pragma unroll
            for (int q = 0; q < 2; q++) { Am[q] = 0; Bm[q] = 1; Cm[q] = 0; }
#pragma unroll
            for (int v = 0; v < 4; v++) Dm[v] = 0;
        }
        if (!hp) {
#pragma unroll
            for (int q = 0; q < 2; q++) { Ap[q] = 0; Bp[q] = 1; Cp[q] = 0; }
#pragma unroll
            for (int v = 0; v < 4; v++) Dp[v] = 0;
        }
#pragma unroll
        for (int q = 0; q < 2; q++) {
            T k1 = A[q] / Bm[q];
            T k2 = C[q] / Bp[q];
            B[q] = B[q] - Cm[q] * k1 - Ap[q] * k2;
            if (q == 0) {
#pragma unroll
                for (int v = 0; v < 3; v++) D[v] = D[v] - Dm[v] * k1 - Dp[v] * k2;
            } else {
                D[3] = D[3] - Dm[3] * k1 - Dp[3] * k2;
            }
            A[q] = -Am[q] * k1;
            C[q] = -Cp[q] * k2;
        }
    }
#pragma unroll
    for (int v = 0; v < 3; v++) D[v] = D[v] / B[0];
    D[3] = D[3] / B[1];
}

// Reduced systems staged in LDS as [line][field][row] with a line pitch of
// kRedFields*R + 1 elements (conflict-free for both the chunk writes, one lane
// per line, and the PCR reads, one lane per row).
struct RedLayout {
    int R, pitch;
    __device__ __forceinline__ int at(int line, int field, int row) const {
        return line * pitch + field * R + row;
    }
};

// Solves `nl` reduced systems held in LDS with the lanes of `nw` wavefronts
// (this one is wave `w`); writes x over the D fields.
template <class T>
__device__ __forceinline__ void solve_reduced_lines(T* red, RedLayout L, int nl, int w, int nw) {
    const int lane = threadIdx.x & 63;
    int RP = 1;
    while (RP < L.R) RP <<= 1;
    const int lpp = 64 / RP;                 // lines per pass of one wave
    const int i = lane % RP;
    for (int base = w * lpp; base < nl; base += nw * lpp) {
        const int line = base + lane / RP;
        const int Rl = line < nl ? L.R : 0;  // rows of this lane's system (0: no system)
        const bool act = i < Rl;
        T A[2], B[2], C[2], D[4];
        if (act) {
            A[0] = red[L.at(line, 0, i)]; B[0] = red[L.at(line, 1, i)]; C[0] = red[L.at(line, 2, i)];
            A[1] = red[L.at(line, 3, i)]; B[1] = red[L.at(line, 4, i)]; C[1] = red[L.at(line, 5, i)];
#pragma unroll
            for (int v = 0; v < 4; v++) D[v] = red[L.at(line, 6 + v, i)];
        } else {
            A[0] = A[1] = C[0] = C[1] = 0; B[0] = B[1] = 1;
            D[0] = D[1] = D[2] = D[3] = 0;
        }
        pcr_solve<T>(A, B, C, D, i, Rl, RP);
        if (act) {
#pragma unroll
            for (int v = 0; v < 4; v++) red[L.at(line, 6 + v, i)] = D[v];
        }
    }
}

// ---------------------------------------------------------------------------
// Strided lines (X or Y sweep).  A work-group owns LPG consecutive z lines and
// every row of them (MAXT = 32 * LPG threads for lines of up to 256 rows in
// chunks of M = 8); MINW is the occupancy the register allocation must allow
// (waves per SIMD), so that two half-width work-groups can share a CU.
template <class T, int AXIS, int M, int LPG, int MAXT, int MINW>
__global__ void __launch_bounds__(MAXT, MINW) fused_strided_kernel(
    const InstDesc<T>* __restrict__ insts, T three_dt, const uint8_t* __restrict__ type,
    const uint8_t* __restrict__ bcv, const uint8_t* __restrict__ bct, const T* __restrict__ gv0,
    const T* __restrict__ gv1, const T* __restrict__ gv2, const T* __restrict__ gv3,
    Layer4<T> in, Layer4<T> tin, Layer4<T> out, Layer4<T> tout, int merges,
    const int32_t* __restrict__ work) {
    extern __shared__ char smem_raw[];
    T* red = reinterpret_cast<T*>(smem_raw);
    const T* gv[4] = {gv0, gv1, gv2, gv3};
    constexpr int GPW = 64 / LPG;  // chunks (lane groups) per wave
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const int gl = lane % LPG;
    const int chunk = wave * GPW + lane / LPG;
    const int P = nw * GPW;
    const RedLayout L{2 * P, kRedFields * 2 * P + 1};

    // work item: (instance, tile = outer * ztiles + z tile) of this work-group
    const int b = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int inst = work[2 * b], tile = work[2 * b + 1];
    const Geom g = insts[inst].g;
    SweepParams<T> p = insts[inst].p[AXIS];
    p.three_dt = three_dt;
    const int n = AXIS == AX_X ? g.nx : g.ny;
    const int64_t s = AXIS == AX_X ? g.sx : g.sy;
    const int ztiles = (g.nz + LPG - 1) / LPG;
    const int outer = tile / ztiles;
    const int z = (tile % ztiles) * LPG + gl;
    const bool zok = z < g.nz;
    const int64_t base =
        insts[inst].off + (AXIS == AX_X ? (int64_t)outer * g.sy : (int64_t)outer * g.sx) + z;
    const int r0 = chunk * M;

    int role[M];
    Chunk<T, M> ch;
    T f[kRedFields], l[kRedFields];
    chunk_phase<T, AXIS, M>(g, p, type, bcv, bct, gv, in, tin, base, s, r0, n, zok, role, ch, f, l);
#pragma unroll
    for (int k = 0; k < kRedFields; k++) {
        red[L.at(gl, k, 2 * chunk)] = f[k];
        red[L.at(gl, k, 2 * chunk + 1)] = l[k];
    }
    __syncthreads();
    solve_reduced_lines<T>(red, L, LPG, wave, nw);
    __syncthreads();
    if (!zok) return;
    T xf[4], xl[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        xf[v] = red[L.at(gl, 6 + v, 2 * chunk)];
        xl[v] = red[L.at(gl, 6 + v, 2 * chunk + 1)];
    }
    write_phase<T, M>(ch, role, base, s, r0, xf, xl, tin, out, tout, merges);
}

// 16-byte vector of consecutive elements.
template <class T>
struct Vec16;
template <>
struct Vec16<float> {
    using type = float4;
    static constexpr int n = 4;
};
template <>
struct Vec16<double> {
    using type = double2;
    static constexpr int n = 2;
};

// r[0..M) <- p[0..M): 16-byte loads when `vec` (all M rows inside the line and
// aligned), else element loads of the first `valid` rows (zero beyond).
template <class T, int M>
__device__ __forceinline__ void load_run(const T* __restrict__ p, T* r, bool vec, int valid) {
    constexpr int VW = Vec16<T>::n;
    if constexpr (M % VW == 0) {
        if (vec) {
            using V = typename Vec16<T>::type;
#pragma unroll
            for (int k = 0; k < M / VW; k++) {
                V v = reinterpret_cast<const V*>(p)[k];
                const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
                for (int q = 0; q < VW; q++) r[k * VW + q] = e[q];
            }
            return;
        }
    }
#pragma unroll
    for (int j = 0; j < M; j++) r[j] = j < valid ? p[j] : (T)0;
}

template <class T, int M>
__device__ __forceinline__ void store_run(T* __restrict__ p, const T* r, bool vec, int valid) {
    constexpr int VW = Vec16<T>::n;
    if constexpr (M % VW == 0) {
        if (vec) {
            using V = typename Vec16<T>::type;
#pragma unroll
            for (int k = 0; k < M / VW; k++) {
                V v;
                T* e = reinterpret_cast<T*>(&v);
#pragma unroll
                for (int q = 0; q < VW; q++) e[q] = r[k * VW + q];
                reinterpret_cast<V*>(p)[k] = v;
            }
            return;
        }
    }
#pragma unroll
    for (int j = 0; j < M; j++)
        if (j < valid) p[j] = r[j];
}

// Contiguous lines (Z sweep): 32 lanes per line, two lines per wavefront.  A
// lane's chunk is contiguous in memory, so it is read and written with 16-byte
// vector accesses (the wave covers its two lines with back-to-back vectors);
// the along-line neighbours of a chunk come from the adjacent lanes by shuffle
// and the cross-line derivatives of w from vector loads of the neighbour lines.
template <class T, int M, int MINW>
__global__ void __launch_bounds__(64 * kZWaves, MINW) fused_z_kernel(
    const InstDesc<T>* __restrict__ insts, T three_dt, const uint8_t* __restrict__ type,
    const uint8_t* __restrict__ bcv, const uint8_t* __restrict__ bct, const T* __restrict__ gv0,
    const T* __restrict__ gv1, const T* __restrict__ gv2, const T* __restrict__ gv3,
    Layer4<T> in, Layer4<T> tin, Layer4<T> out, Layer4<T> tout, int merges,
    const int32_t* __restrict__ lines, const int32_t* __restrict__ winst, int vec_ok) {
    extern __shared__ char smem_raw[];
    const T* gv[4] = {gv0, gv1, gv2, gv3};
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    // every line of a work-group belongs to one instance
    const int b = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int inst = winst[b];
    const Geom g = insts[inst].g;
    SweepParams<T> p = insts[inst].p[AX_Z];
    p.three_dt = three_dt;
    const int n = g.nz;
    const int used = (n + M - 1) / M;  // lanes of a line holding rows (<= 32)
    const RedLayout L{2 * used, kRedFields * 2 * used + 1};
    T* red = reinterpret_cast<T*>(smem_raw) + (int64_t)wave * 2 * L.pitch;

    const int sub = lane >> 5, ll = lane & 31;
    const int lid = lines[b * (2 * nw) + 2 * wave + sub];  // x*ny + y, or -1 (padding)
    const bool live = lid >= 0;
    const int64_t base = live ? insts[inst].off + (int64_t)lid * g.nz : 0;
    const int r0 = ll * M;
    const bool ok = live && ll < used;
    const int valid = ok ? min(M, n - r0) : 0;
    const bool vec = vec_ok && valid == M;
    const int64_t c0 = base + r0;

    // node types of the chunk and of the rows either side (from the neighbour lanes)
    uint8_t ty[M + 2];
#pragma unroll
    for (int j = 0; j < M; j++) ty[j + 1] = j < valid ? type[c0 + j] : kOutside;
    {
        int prev = __shfl((int)ty[M], lane - 1, 64);
        int next = __shfl((int)ty[1], lane + 1, 64);
        ty[0] = (ll > 0 && ok) ? (uint8_t)prev : kOutside;
        ty[M + 1] = (ll < 31 && ok && r0 + M < n) ? (uint8_t)next : kOutside;
    }
    int role[M];
#pragma unroll
    for (int j = 0; j < M; j++) role[j] = row_role(ty[j + 1], ty[j], ty[j + 2]);

    // inputs: state (in), Picard layer (tin) and the cross-line derivatives of w
    T cu[4][M], tv[4][M], wx[M], wy[M];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        load_run<T, M>(in.f[v] + c0, cu[v], vec, valid);
        load_run<T, M>(tin.f[v] + c0, tv[v], vec, valid);
    }
    {
        // the active lines have x, y in [1, n-2], so the neighbour lines exist
        T a[M], b[M];
        const T* tw = tin.f[2];
        load_run<T, M>(tw + c0 + (ok ? g.sx : 0), a, vec, valid);
        load_run<T, M>(tw + c0 - (ok ? g.sx : 0), b, vec, valid);
#pragma unroll
        for (int j = 0; j < M; j++) wx[j] = (a[j] - b[j]) * p.inv_2hx;
        load_run<T, M>(tw + c0 + (ok ? g.sy : 0), a, vec, valid);
        load_run<T, M>(tw + c0 - (ok ? g.sy : 0), b, vec, valid);
#pragma unroll
        for (int j = 0; j < M; j++) wy[j] = (a[j] - b[j]) * p.inv_2hy;
    }
    T tprev[4], tnext[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        tprev[v] = __shfl(tv[v][M - 1], lane - 1, 64);
        tnext[v] = __shfl(tv[v][0], lane + 1, 64);
    }

    Chunk<T, M> ch;
    Row<T> rfirst, rlast;
#pragma unroll
    for (int j = 0; j < M; j++) {
        Row<T> r;
        if (role[j] == RL_IN) {
            const T adv = tv[2][j] * p.inv_2h;
            r.av = -adv - p.vis_v;
            r.cv = adv - p.vis_v;
            r.bv = p.three_dt + 2 * p.vis_v;
            r.at = -adv - p.vis_t;
            r.ct = adv - p.vis_t;
            r.bt = p.three_dt + 2 * p.vis_t;
            T m1[4], p1[4];
#pragma unroll
            for (int v = 0; v < 4; v++) {
                m1[v] = j > 0 ? tv[v][j - 1] : tprev[v];
                p1[v] = j < M - 1 ? tv[v][j + 1] : tnext[v];
                r.d[v] = cu[v][j] * p.three_dt;
            }
            r.d[2] = r.d[2] - p.v_T * ((p1[3] - m1[3]) * p.inv_2h);
            const T u_z = (p1[0] - m1[0]) * p.inv_2hz;
            const T v_z = (p1[1] - m1[1]) * p.inv_2hz;
            const T w_z = (p1[2] - m1[2]) * p.inv_2hz;
            const T phi = u_z * u_z + v_z * v_z + 2 * w_z * w_z + u_z * wx[j] + v_z * wy[j];
            r.d[3] = r.d[3] + p.t_phi * phi;
        } else {
            boundary_row<T>(role[j], c0 + j, bcv, bct, gv, r);
        }
        if (j == 0) rfirst = r;
        else if (j == M - 1) rlast = r;
        else chunk_down<T, M>(ch, j, r);
    }
    if (M >= 3) chunk_up<T, M>(ch);
    if (ll < used) {
        T f[kRedFields], l[kRedFields];
        chunk_reduced<T, M>(ch, rfirst, rlast, f, l);
#pragma unroll
        for (int k = 0; k < kRedFields; k++) {
            red[L.at(sub, k, 2 * ll)] = f[k];
            red[L.at(sub, k, 2 * ll + 1)] = l[k];
        }
    }
    __syncthreads();
    solve_reduced_lines<T>(red, L, 2, 0, 1);
    __syncthreads();
    if (!ok) return;
    bool all_written = true;
#pragma unroll
    for (int j = 0; j < M; j++) all_written = all_written && role[j] != RL_NONE;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const T xf = red[L.at(sub, 6 + v, 2 * ll)];
        const T xl = red[L.at(sub, 6 + v, 2 * ll + 1)];
        T x[M], t[M];
#pragma unroll
        for (int j = 0; j < M; j++) {
            if (j == 0) x[j] = xf;
            else if (j == M - 1) x[j] = xl;
            else {
                const T F = v < 3 ? ch.Fv[j] : ch.Ft[j];
                const T G = v < 3 ? ch.Gv[j] : ch.Gt[j];
                x[j] = ch.E[v][j] + F * xf + G * xl;
            }
            // the Picard layer keeps its value off NODE_IN (rewritten unchanged)
            T m = (tv[v][j] + x[j]) * (T)0.5;
            if (merges > 1) m = (m + x[j]) * (T)0.5;
            t[j] = role[j] == RL_IN ? m : tv[v][j];
        }
        if (all_written && vec) {
            store_run<T, M>(out.f[v] + c0, x, true, M);
        } else {
#pragma unroll
            for (int j = 0; j < M; j++)
                if (j < valid && role[j] != RL_NONE) out.f[v][c0 + j] = x[j];
        }
        store_run<T, M>(tout.f[v] + c0, t, vec, valid);
    }
}

constexpr int kChunk = 8;  // rows per chunk of the strided kernel

// Lines per work-group of a strided-sweep variant: 0 = full 128-byte rows
// (32 * LPG threads, one work-group per CU), 1 = half rows of 64 bytes with
// registers capped for two work-groups per CU (two work-groups share each
// cache line).
template <class T>
constexpr int variant_lpg(int variant) { return (variant == 0 ? 128 : 64) / (int)sizeof(T); }

int strided_waves_for(int n, int lpg) {
    const int gpw = 64 / lpg;
    return (n + kChunk * gpw - 1) / (kChunk * gpw);
}

int z_chunk(int n) {
    for (int m : {2, 4, 8, 16})
        if (32 * m >= n) return m;
    return 0;
}

}  // namespace

template <class T>
int strided_lpg(int variant) {
    return variant_lpg<T>(variant);
}

template <class T>
bool fused_supported(const Geom& g) {
    if (z_chunk(g.nz) == 0 || (sizeof(T) == 8 && z_chunk(g.nz) > 8)) return false;
    // both variants: at most 32 * LPG threads
    for (int v = 0; v < 2; v++) {
        const int lpg = variant_lpg<T>(v);
        if (strided_waves_for(g.nx, lpg) * 64 > 32 * lpg) return false;
        if (strided_waves_for(g.ny, lpg) * 64 > 32 * lpg) return false;
    }
    return true;
}

template <class T>
void launch_fused_sweep(const InstDesc<T>* insts, T three_dt, int axis, const FusedWork& w,
                        const uint8_t* type, const uint8_t* bcv, const uint8_t* bct,
                        const T* const* gval, Layer4<T> in, Layer4<T> tin, Layer4<T> out,
                        Layer4<T> tout, int merges, int variant, hipStream_t st) {
    if (w.nblocks == 0) return;
    const unsigned blocks = (unsigned)w.nblocks;
    if (axis == AX_Z) {
        const int m = z_chunk(w.maxn);
        const int used = (w.maxn + m - 1) / m;
        const int pitch = kRedFields * 2 * used + 1;
        size_t lds = (size_t)kZWaves * 2 * pitch * sizeof(T);
        // variant 1 caps the registers for 4 (FP32) / 2 (FP64) waves per SIMD
        constexpr int MW = sizeof(T) == 8 ? 2 : 4;
#define LZ(MM, W)                                                                               \
    hipLaunchKernelGGL((fused_z_kernel<T, MM, W>), dim3(blocks), dim3(64 * kZWaves), lds, st,  \
                       insts, three_dt, type, bcv, bct, gval[0], gval[1], gval[2], gval[3], in, \
                       tin, out, tout, merges, w.d_items, w.d_inst, w.vec_ok)
        if (variant == 0) {
            if (m == 2) LZ(2, 1);
            else if (m == 4) LZ(4, 1);
            else if (m == 8) LZ(8, 1);
            else LZ(16, 1);
        } else {
            if (m == 2) LZ(2, MW);
            else if (m == 4) LZ(4, MW);
            else if (m == 8) LZ(8, MW);
            else LZ(16, 1);
        }
#undef LZ
    } else {
        constexpr int LW = variant_lpg<T>(0), LN = variant_lpg<T>(1);
        const int lpg = variant ? LN : LW;
        const int waves = strided_waves_for(w.maxn, lpg);
        const int P = waves * (64 / lpg);
        const int pitch = kRedFields * 2 * P + 1;
        size_t lds = (size_t)lpg * pitch * sizeof(T);
#define LS(AX, LPGV)                                                                             \
    hipLaunchKernelGGL((fused_strided_kernel<T, AX, kChunk, LPGV, 32 * LPGV,                      \
                                             (LPGV == LN ? LPGV / 4 : LPGV / 8)>),                \
                       dim3(blocks), dim3(64 * waves), lds, st, insts, three_dt, type, bcv, bct,  \
                       gval[0], gval[1], gval[2], gval[3], in, tin, out, tout, merges, w.d_items)
        if (axis == AX_X) {
            if (variant) LS(AX_X, LN);
            else LS(AX_X, LW);
        } else {
            if (variant) LS(AX_Y, LN);
            else LS(AX_Y, LW);
        }
#undef LS
    }
    CMCX_LAUNCH_CHECK();
}

void FusedWork::release() {
    hipFree(d_items);
    hipFree(d_inst);
    d_items = d_inst = nullptr;
    nblocks = 0;
}

// Work of one axis over all instances.  Z: lines x*ny + y holding a NODE_IN
// node, each instance's list padded with -1 to whole work-groups (2 lines per
// wave), plus the instance of every work-group.  X/Y: (instance, tile) pairs,
// tile = outer * ztiles + z tile, for the tiles of lpg z lines holding a
// NODE_IN node.
void build_fused_work(const uint8_t* h_type, const std::vector<Geom>& geoms,
                      const std::vector<int64_t>& offs, int axis, int lpg, size_t elem,
                      FusedWork& w) {
    w.release();
    std::vector<int32_t> items, winst;
    w.maxn = 0;
    w.vec_ok = 1;
    for (size_t i = 0; i < geoms.size(); i++) {
        const Geom& g = geoms[i];
        const uint8_t* t = h_type + offs[i];
        if (axis == AX_Z) {
            w.maxn = std::max(w.maxn, g.nz);
            if (((size_t)g.nz * elem) % 16 || ((size_t)offs[i] * elem) % 16) w.vec_ok = 0;
            size_t first = items.size();
            for (int x = 0; x < g.nx; x++)
                for (int y = 0; y < g.ny; y++) {
                    const uint8_t* row = &t[x * g.sx + y * g.sy];
                    bool any = false;
                    for (int z = 0; z < g.nz && !any; z++) any = row[z] == NODE_IN;
                    if (any) items.push_back(x * g.ny + y);
                }
            const size_t per = 2 * kZWaves;
            while ((items.size() - first) % per) items.push_back(-1);
            for (size_t k = first; k < items.size(); k += per) winst.push_back((int32_t)i);
        } else {
            const int n = axis == AX_X ? g.nx : g.ny;
            w.maxn = std::max(w.maxn, n);
            const int ztiles = (g.nz + lpg - 1) / lpg;
            const int nouter = axis == AX_X ? g.ny : g.nx;
            const int64_t so = axis == AX_X ? g.sy : g.sx, sr = axis == AX_X ? g.sx : g.sy;
            for (int o = 0; o < nouter; o++)
                for (int zt = 0; zt < ztiles; zt++) {
                    bool any = false;
                    for (int r = 0; r < n && !any; r++)
                        for (int z = zt * lpg; z < std::min(g.nz, (zt + 1) * lpg) && !any; z++)
                            any = t[o * so + r * sr + z] == NODE_IN;
                    if (any) {
                        items.push_back((int32_t)i);
                        items.push_back(o * ztiles + zt);
                    }
                }
        }
    }
    w.nblocks = axis == AX_Z ? (int64_t)winst.size() : (int64_t)items.size() / 2;
    if (!items.empty()) {
        CMCX_HIP(hipMalloc((void**)&w.d_items, sizeof(int32_t) * items.size()));
        CMCX_HIP(hipMemcpy(w.d_items, items.data(), sizeof(int32_t) * items.size(),
                           hipMemcpyHostToDevice));
    }
    if (!winst.empty()) {
        CMCX_HIP(hipMalloc((void**)&w.d_inst, sizeof(int32_t) * winst.size()));
        CMCX_HIP(hipMemcpy(w.d_inst, winst.data(), sizeof(int32_t) * winst.size(),
                           hipMemcpyHostToDevice));
    }
}

int shared_free_axes(const uint8_t* t, const uint8_t* bv, const uint8_t* bt, const Geom& g) {
    int mask = 0;
    const int64_t str[3] = {g.sx, g.sy, 1};
    const int dim[3] = {g.nx, g.ny, g.nz};
    for (int x = 0; x < g.nx; x++)
        for (int y = 0; y < g.ny; y++)
            for (int z = 0; z < g.nz; z++) {
                int64_t c = x * g.sx + y * g.sy + z;
                if (t[c] == NODE_IN || (bv[c] != BC_FREE && bt[c] != BC_FREE)) continue;
                const int pos[3] = {x, y, z};
                for (int a = 0; a < 3; a++)
                    if (pos[a] > 0 && pos[a] < dim[a] - 1 && t[c - str[a]] == NODE_IN &&
                        t[c + str[a]] == NODE_IN)
                        mask |= 1 << a;
            }
    return mask;
}

template bool fused_supported<float>(const Geom&);
template bool fused_supported<double>(const Geom&);
template int strided_lpg<float>(int);
template int strided_lpg<double>(int);
template void launch_fused_sweep<float>(const InstDesc<float>*, float, int, const FusedWork&,
                                        const uint8_t*, const uint8_t*, const uint8_t*,
                                        const float* const*, Layer4<float>, Layer4<float>,
                                        Layer4<float>, Layer4<float>, int, int, hipStream_t);
template void launch_fused_sweep<double>(const InstDesc<double>*, double, int, const FusedWork&,
                                         const uint8_t*, const uint8_t*, const uint8_t*,
                                         const double* const*, Layer4<double>, Layer4<double>,
                                         Layer4<double>, Layer4<double>, int, int, hipStream_t);

}  // namespace cmcx
