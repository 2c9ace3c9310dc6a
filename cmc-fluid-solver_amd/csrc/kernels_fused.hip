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
//   otherwise               identity row x = 0, not stored (Z: see "Stores")
// On a well-formed mask the line-wide matrix is block diagonal with exactly the
// reference's segments (Grid3D.cpp:62-118) as blocks, so solving it gives every
// segment's Thomas solution up to rounding.  Outputs follow UpdateSegment
// (every segment node written) and the Picard merge on NODE_IN nodes.
//
// Stores: the Z sweep writes every row of a line with whole-vector stores (a
// chunk stored with gaps costs far more than its bytes); rows no segment covers
// receive the sweep input's value.  That is what the reference leaves in its
// output `next` at the end of the step: every layer starts as the state, the
// step set-up copies BOUND and VALVE values of the state into next, and where
// the X sweep that follows covers the node it overwrites it, here too.  The
// strided (X, Y) sweeps store segment rows only, as the reference (storing the
// other rows as well costs them more in bytes than it gains).  When a launch's
// 16 fields exceed the memory-side cache, all stores are non-temporal
// (kStoreStream): written lines would otherwise evict lines that the sweep
// still reads (the Picard layer for the merge, the cross-line rows).
//
// Each line is solved by the chunked partition method of line_solve.h (chunks
// of M rows per lane, reduced system of 2 rows per chunk, compacted by one
// cyclic-reduction step to one row per chunk, at most 32 rows per line, then
// PCR across the lanes: through LDS for strided lines, in registers for Z).  The U, V and W systems share one matrix (same
// bc_vel and viscosity); T has its own.
//
// Node flags: the fused kernels read one byte per node, the node type in bits
// 0-1 and the BC kinds (BC_FREE) of velocity and T in bits 2 and 3 (0 for
// NODE_IN).  Dirichlet values of boundary rows come from the Picard layer
// while the state is in sync with the node values (after UpdateBoundaries;
// steps keep it: their no-slip rows write the node values), else from the
// node-value arrays (`bsync`).
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
#include "line_solve.h"

#include <algorithm>
#include <numeric>
#include <stdexcept>

namespace cmcx {

namespace {

using namespace lsolve;

// Correction form.  Every sweep solves for the correction delta = x - t to
// the Picard layer t (the current iterate) instead of for x itself: A delta =
// d - A t, with t on every row of a system.  For an interior row (b = 3/dt +
// 2 nu/h^2, a + c = -2 nu/h^2) the right-hand side is
//   3/dt (x_in - t_j) + adv (t_{j-1} - t_{j+1}) + nu/h^2 (t_{j-1} + t_{j+1} - 2 t_j) + e_j,
// e_j the pressure term or v_T-scaled T gradient of the field along the sweep,
// or t_phi Phi_d for T: the same discrete equations as the reference's rows
// (AdiSolver3D.cpp:755-801), with every term a difference of neighbouring
// values.  Solving for x directly rounds the right-hand side 3/dt x_in and
// the whole solution at the scale of x, and in FP32 that hides the small
// components: v and w beside u, and T - baseT beside T ~ baseT, which the
// correction keeps at their own scale.  Boundary rows: a Dirichlet row takes
// delta = value - t (0 while the state is in sync: the Picard layer holds
// the node value); a free row (2 x_j - x_{j+-1} = 0) takes t_{j+-1} - 2 t_j.
// The sweep writes x = t + delta (the value itself on Dirichlet rows) and the
// merged Picard layer t + delta/2 (t + 3 delta/4 after the folded second
// merge).

// Interior row in correction form: wm, w0, wp are the Picard layer at the rows
// before, at and after the node (per field), cin the sweep input at the node,
// ca and cb the central differences (not yet scaled) of the cross-line
// derivatives of Phi_d (X: u along y and z; Y: v along x and z; Z: w along x
// and y) (TimeLayer3D.h:338-343, 554-588).
template <class T, int AXIS>
__device__ __forceinline__ void interior_row_delta(const SweepParams<T>& p, const T* wm,
                                                   const T* w0, const T* wp, const T* cin, T ca,
                                                   T cb, Row<T>& r) {
    const T adv = w0[AXIS] * p.inv_2h;
    r.av = -adv - p.vis_v;
    r.cv = adv - p.vis_v;
    r.bv = p.three_dt + 2 * p.vis_v;
    r.at = -adv - p.vis_t;
    r.ct = adv - p.vis_t;
    r.bt = p.three_dt + 2 * p.vis_t;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const T vis = v < 3 ? p.vis_v : p.vis_t;
        r.d[v] = p.three_dt * (cin[v] - w0[v]) + adv * (wm[v] - wp[v]) +
                 vis * ((wm[v] + wp[v]) - 2 * w0[v]);
    }
    r.d[AXIS] = r.d[AXIS] - p.v_T * ((wp[3] - wm[3]) * p.inv_2h);
    T phi;
    if (AXIS == AX_X) {
        const T u_x = (wp[0] - wm[0]) * p.inv_2hx, v_x = (wp[1] - wm[1]) * p.inv_2hx;
        const T w_x = (wp[2] - wm[2]) * p.inv_2hx;
        const T u_y = ca * p.inv_2hy, u_z = cb * p.inv_2hz;
        phi = 2 * u_x * u_x + v_x * v_x + w_x * w_x + v_x * u_y + w_x * u_z;
    } else if (AXIS == AX_Y) {
        const T u_y = (wp[0] - wm[0]) * p.inv_2hy, v_y = (wp[1] - wm[1]) * p.inv_2hy;
        const T w_y = (wp[2] - wm[2]) * p.inv_2hy;
        const T v_x = ca * p.inv_2hx, v_z = cb * p.inv_2hz;
        phi = u_y * u_y + 2 * v_y * v_y + w_y * w_y + u_y * v_x + w_y * v_z;
    } else {
        const T u_z = (wp[0] - wm[0]) * p.inv_2hz, v_z = (wp[1] - wm[1]) * p.inv_2hz;
        const T w_z = (wp[2] - wm[2]) * p.inv_2hz;
        const T w_x = ca * p.inv_2hx, w_y = cb * p.inv_2hy;
        phi = u_z * u_z + v_z * v_z + 2 * w_z * w_z + u_z * w_x + v_z * w_y;
    }
    r.d[3] = r.d[3] + p.t_phi * phi;
}

// Boundary or identity row (ApplyBC0 / ApplyBC1, AdiSolver3D.cpp:804-852) in
// correction form, without branches: tf is the node's flags byte (BC kinds in
// bits 2 and 3), gv the Dirichlet values.
template <class T>
__device__ __forceinline__ void boundary_row_delta(int role, uint8_t tf, const T* wm, const T* w0,
                                                   const T* wp, const T* gv, Row<T>& r) {
    const bool fsv = role == RL_START && (tf & kFlagFreeVel), fev = role == RL_END && (tf & kFlagFreeVel);
    const bool fst = role == RL_START && (tf & kFlagFreeT), fet = role == RL_END && (tf & kFlagFreeT);
    const bool val = role != RL_NONE;
    r.av = fev ? (T)-1 : (T)0;
    r.bv = (fsv || fev) ? (T)2 : (T)1;
    r.cv = fsv ? (T)-1 : (T)0;
    r.at = fet ? (T)-1 : (T)0;
    r.bt = (fst || fet) ? (T)2 : (T)1;
    r.ct = fst ? (T)-1 : (T)0;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const bool fs = v < 3 ? fsv : fst, fe = v < 3 ? fev : fet;
        const T dir = val ? gv[v] - w0[v] : (T)0;
        r.d[v] = fs ? wp[v] - 2 * w0[v] : fe ? wm[v] - 2 * w0[v] : dir;
    }
}

// Whether row `role` of matrix q (0 velocity, 1 T) is a Dirichlet row.
__device__ __forceinline__ bool dirichlet_row(int role, uint8_t tf, int q) {
    const uint8_t fb = q == 0 ? kFlagFreeVel : kFlagFreeT;
    return role != RL_NONE && role != RL_IN && !((role == RL_START || role == RL_END) && (tf & fb));
}

// Strided chunk: loads the chunk's rows (and the Picard layer one row either
// side), eliminates them and returns the two reduced rows.  Row r of the line
// lives at element base + r * s; rows outside [0, n), and every row when !ok,
// read the interior node `safe` instead (so that each load is unconditional
// and in bounds) and are identity rows.  The loads are software-pipelined:
// those of row j + LA are issued before row j is eliminated, in one basic
// block, so that LA + 1 rows are in flight instead of one (a row whose loads
// sit in a branch of its own waits for all of them before the next row's
// loads can issue).  role[] receives the row roles.  KEEP parks the chunk's
// Picard layer in LDS for the write phase.
template <class T, int AXIS, int M, bool KEEP, bool BSYNC>
__device__ __forceinline__ void strided_chunk(const Geom& g, const SweepParams<T>& p,
                                              const uint8_t* __restrict__ type,
                                              const T* const* gval, const Layer4<T>& in,
                                              const Layer4<T>& tin, int base, int s, int safe,
                                              int r0, int n, bool ok, int* role, Chunk<T, M>& ch,
                                              T* f, T* l, T* keep, int kfs, int krs) {
    constexpr int LA = 2;  // rows loaded ahead of the elimination
    auto addr = [&](int i) {  // element of window row i (line row r0 - 1 + i)
        const int r = r0 + i - 1;
        return (ok && r >= 0 && r < n) ? base + r * s : safe;
    };
    auto valid = [&](int i) {
        const int r = r0 + i - 1;
        return ok && r >= 0 && r < n;
    };
    uint8_t ty[M + 2];
#pragma unroll
    for (int i = 0; i < M + 2; i++) {
        const uint8_t t = type[addr(i)];
        ty[i] = valid(i) ? t : kOutside;
    }
    constexpr int XF = AXIS == AX_X ? 0 : 1;  // field of the cross-line derivatives
    const int sa = AXIS == AX_X ? (int)g.sy : (int)g.sx;
    const T* xfp = tin.f[XF];
    // Picard layer window (rows r0-1 .. r0+M), sweep input and cross-line
    // differences, filled LA rows ahead of their use
    T w[4][M + 2], cin[4][M], ca[M], cb[M];
    auto load_row = [&](int j) {  // window row j + 1 and row j's own values
        const int c = addr(j + 1);
#pragma unroll
        for (int v = 0; v < 4; v++) w[v][j + 1] = tin.f[v][c];
#pragma unroll
        for (int v = 0; v < 4; v++) cin[v][j] = in.f[v][c];
        ca[j] = xfp[c + sa] - xfp[c - sa];
        cb[j] = xfp[c + 1] - xfp[c - 1];
    };
    {
        const int c = addr(0);
#pragma unroll
        for (int v = 0; v < 4; v++) w[v][0] = tin.f[v][c];
    }
#pragma unroll
    for (int j = 0; j < LA && j < M; j++) load_row(j);
#pragma unroll
    for (int j = 0; j < M; j++) role[j] = row_role(ty[j + 1], ty[j], ty[j + 2]);
    Row<T> rfirst, rlast;
#pragma unroll
    for (int j = 0; j < M; j++) {
        if (j + LA < M) load_row(j + LA);
        if (j + LA == M - 1 + 1 || (j == M - 1 && LA == 0)) {
            const int c = addr(M + 1);
#pragma unroll
            for (int v = 0; v < 4; v++) w[v][M + 1] = tin.f[v][c];
        }
        T wm[4], w0[4], wp[4], ci[4], gd[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            wm[v] = w[v][j];
            w0[v] = w[v][j + 1];
            wp[v] = w[v][j + 2];
            ci[v] = cin[v][j];
            gd[v] = BSYNC ? w0[v] : gval[v][addr(j + 1)];
        }
        if (KEEP) {
#pragma unroll
            for (int v = 0; v < 4; v++) keep[v * kfs + j * krs] = w0[v];
        }
        Row<T> ri, rb, rw;
        interior_row_delta<T, AXIS>(p, wm, w0, wp, ci, ca[j], cb[j], ri);
        boundary_row_delta<T>(role[j], ty[j + 1], wm, w0, wp, gd, rb);
        const bool isin = role[j] == RL_IN;
        rw.av = isin ? ri.av : rb.av; rw.bv = isin ? ri.bv : rb.bv; rw.cv = isin ? ri.cv : rb.cv;
        rw.at = isin ? ri.at : rb.at; rw.bt = isin ? ri.bt : rb.bt; rw.ct = isin ? ri.ct : rb.ct;
#pragma unroll
        for (int v = 0; v < 4; v++) rw.d[v] = isin ? ri.d[v] : rb.d[v];
        if (j == 0) rfirst = rw;
        else if (j == M - 1) rlast = rw;
        else chunk_down<T, M>(ch, j, rw);
        // keep the pipeline's shape: the scheduler would otherwise hoist the
        // loads of every row to the top and spill what does not fit
        __builtin_amdgcn_sched_barrier(0);
    }
    if (M >= 3) chunk_up<T, M>(ch);
    chunk_reduced<T, M>(ch, rfirst, rlast, f, l);
}

// Writes the chunk's rows given the corrections df / dl of its first and last
// rows: x = t + delta on every row of a system (the node value on Dirichlet
// rows), the merged Picard layer on NODE_IN rows.  Row j lives at element
// base + (r0 + j) * s.  The corrections of all rows are formed first (the
// chunk state dies there), then the Picard layer of all rows is read, then
// everything is stored: a load issued after a store waits for the store's
// acknowledgement as well (one counter covers both on this architecture), so
// reads interleaved with the stores would cost a memory round trip per field.
template <class T, int M, bool KEEP, bool BSYNC>
__device__ __forceinline__ void strided_write(const Chunk<T, M>& ch, const int* role,
                                              const uint8_t* __restrict__ type, int base, int s,
                                              int safe, int r0, const T* df, const T* dl,
                                              const T* const* gval, const Layer4<T>& tin,
                                              const Layer4<T>& out, const Layer4<T>& tout,
                                              int merges, int stores, const T* keep, int kfs,
                                              int krs) {
    // the row addresses are formed here, not kept live from the kernel's start
    // (where the compiler would hoist them and spill them across the solve)
    asm volatile("" : "+v"(base), "+v"(r0), "+v"(safe));
    const T km = merges > 1 ? (T)0.75 : (T)0.5;
    T dd[4][M];
#pragma unroll
    for (int j = 0; j < M; j++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            if (j == 0) dd[v][j] = df[v];
            else if (j == M - 1) dd[v][j] = dl[v];
            else {
                const T F = v < 3 ? ch.Fv[j] : ch.Ft[j];
                const T G = v < 3 ? ch.Gv[j] : ch.Gt[j];
                dd[v][j] = ch.E[v][j] + F * df[v] + G * dl[v];
            }
        }
    T t0[4][M];
#pragma unroll
    for (int j = 0; j < M; j++) {
        const int c = role[j] != RL_NONE ? base + (r0 + j) * s : safe;
#pragma unroll
        for (int v = 0; v < 4; v++) t0[v][j] = KEEP ? keep[v * kfs + j * krs] : tin.f[v][c];
    }
    // every read above completes here, before the first store (otherwise the
    // compiler sinks each read into the branch of its row's stores)
#pragma unroll
    for (int j = 0; j < M; j++)
#pragma unroll
        for (int v = 0; v < 4; v++) asm volatile("" ::"v"(t0[v][j]));
    with_store_hint((stores & kStoreStream) != 0, [&](auto ntc) {
        constexpr bool NT = decltype(ntc)::value;
#pragma unroll
        for (int j = 0; j < M; j++) {
            if (role[j] == RL_NONE) continue;
            const int c = base + (r0 + j) * s;
            const uint8_t tf = BSYNC ? (uint8_t)0 : type[c];
#pragma unroll
            for (int v = 0; v < 4; v++) {
                T x = t0[v][j] + dd[v][j];
                if (!BSYNC && dirichlet_row(role[j], tf, v < 3 ? 0 : 1)) x = gval[v][c];
                if (stores & 1) put_hint<NT>(&out.f[v][c], x);
                if ((stores & 2) && role[j] == RL_IN) put_hint<NT>(&tout.f[v][c], t0[v][j] + km * dd[v][j]);
            }
        }
    });
}

// Row pitch (elements) of the LDS copy of the Picard layer: LPG lines plus
// padding so that the lane groups of a wavefront (M rows apart) start on
// different banks.
__host__ __device__ constexpr int keep_pitch(int lpg, int m) { return lpg + (lpg >= m ? lpg / m : 0); }
// 128-byte tiles leave no LDS for the padding (their parked layer and reduced
// systems fill the 160 KB exactly)
template <class T>
__host__ __device__ constexpr int keep_row_pitch(int lpg, int m) {
    return lpg * (int)sizeof(T) >= 128 ? lpg : keep_pitch(lpg, m);
}

// ---------------------------------------------------------------------------
// Strided lines (X or Y sweep).  A work-group owns LPG consecutive z lines and
// every row of them (MAXT = 32 * LPG threads for lines of up to 256 rows in
// chunks of M = 8); MINW is the occupancy the register allocation must allow
// (waves per SIMD), so that two half-width work-groups can share a CU.
template <class T, int AXIS, int M, int LPG, bool KEEP, bool BSYNC>
__device__ __forceinline__ void strided_tile(const InstDesc<T>* __restrict__ insts, T three_dt,
                                             const uint8_t* __restrict__ type, const T* const* gval,
                                             const Layer4<T>& in, const Layer4<T>& tin,
                                             const Layer4<T>& out, const Layer4<T>& tout,
                                             int merges, int stores,
                                             const int32_t* __restrict__ work, int kmax) {
    extern __shared__ char smem_raw[];
    T* red = reinterpret_cast<T*>(smem_raw);
    constexpr int GPW = 64 / LPG;  // chunks (lane groups) per wave
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const int gl = lane % LPG;
    const int chunk = wave * GPW + lane / LPG;
    const int P = nw * GPW;
    const RedLayout L{2 * P, kRedFields * 2 * P + 1};

    // work item: the instance and up to kmax of its tiles (tile = outer *
    // ztiles + z tile, -1 = none).  The work-group is sized for the longest
    // line of the batch; a grid with shorter lines packs several tiles into
    // it, tile k on chunks [k * cpl, (k + 1) * cpl) of each z-line slot.  The
    // lines laid end to end form one block-diagonal system (a line starts and
    // ends on a non-IN face node, whose row does not reach across), so the
    // reduced solve treats them as one line.
    const int b = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int32_t* item = work + (int64_t)b * (kmax + 1);
    const int inst = item[0];
    const Geom g = insts[inst].g;
    SweepParams<T> p = insts[inst].p[AXIS];
    p.three_dt = three_dt;
    const int n = AXIS == AX_X ? g.nx : g.ny;
    const int s = AXIS == AX_X ? (int)g.sx : (int)g.sy;
    const int cpl = (n + M - 1) / M;  // chunks per line of this grid
    const int k = chunk / cpl;
    const int tile = k < kmax ? item[1 + k] : -1;
    const int ztiles = (g.nz + LPG - 1) / LPG;
    const int outer = tile >= 0 ? tile / ztiles : 0;
    const int z = (tile >= 0 ? tile % ztiles : 0) * LPG + gl;
    const bool zok = tile >= 0 && z < g.nz;
    // a grid of a batch holds fewer than 2^31 nodes with its offset (checked at
    // creation): 32-bit element indices
    const int off = (int)insts[inst].off;
    const int base = off + (AXIS == AX_X ? outer * (int)g.sy : outer * (int)g.sx) + z;
    // interior node (1, 1, 1): the element read by rows outside the line
    const int safe = off + (int)g.sx + (int)g.sy + 1;
    const int r0 = (chunk - k * cpl) * M;  // first row of the chunk in its line

    // KEEP: each lane parks the Picard layer of its rows in LDS, after the
    // reduced systems, [field][row][line], with a padded line pitch (the lane
    // groups of a wave then hit different banks) where the LDS has room for
    // it; only the lane itself reads it.  The reduced systems then take the
    // 8-value normalised form (kNormFields), which leaves 128-byte tiles room
    // for the parked layer.
    const int red_elems = KEEP ? LPG * kNormFields * pow2_at_least(P) : LPG * L.pitch;
    const int krs = keep_row_pitch<T>(LPG, M);
    const int kfs = P * M * krs;
    T* keep = red + red_elems + chunk * M * krs + gl;
    int role[M];
    Chunk<T, M> ch;
    T f[kRedFields], l[kRedFields];
    strided_chunk<T, AXIS, M, KEEP, BSYNC>(g, p, type, gval, in, tin, base, s, safe, r0, n, zok,
                                           role, ch, f, l, keep, kfs, krs);
    T xf[4], xl[4];
    if constexpr (KEEP) {
        const NormLayout NL{pow2_at_least(P)};
        T fh[kNormFields];
        normalise_first<T>(f, fh);
#pragma unroll
        for (int k = 0; k < kNormFields; k++) red[NL.at(gl, k, chunk)] = fh[k];
        __syncthreads();
        {
            T nx[kNormFields], cr[kNormFields];
            const bool hn = chunk + 1 < P;
#pragma unroll
            for (int k = 0; k < kNormFields; k++) nx[k] = hn ? red[NL.at(gl, k, chunk + 1)] : (T)0;
            compact_norm<T>(fh, l, nx, cr);
            __syncthreads();  // every lane has read the first rows
#pragma unroll
            for (int k = 0; k < kNormFields; k++) red[NL.at(gl, k, chunk)] = cr[k];
        }
        __syncthreads();
        solve_norm_lines<T>(red, NL, LPG, P, wave, nw);
        __syncthreads();
        if (!zok) return;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int q = v < 3 ? 0 : 1;
            xl[v] = red[NL.at(gl, 4 + v, chunk)];
            const T xlp = chunk > 0 ? red[NL.at(gl, 4 + v, chunk - 1)] : (T)0;
            xf[v] = fh[4 + v] - fh[2 * q] * xlp - fh[2 * q + 1] * xl[v];
        }
    } else {
#pragma unroll
        for (int k = 0; k < kRedFields; k++) {
            red[L.at(gl, k, 2 * chunk)] = f[k];
            red[L.at(gl, k, 2 * chunk + 1)] = l[k];
        }
        __syncthreads();
        solve_reduced_compact<T>(red, L, LPG, wave, nw);
        __syncthreads();
        if (!zok) return;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            xf[v] = red[L.at(gl, 6 + v, 2 * chunk)];
            xl[v] = red[L.at(gl, 6 + v, 2 * chunk + 1)];
        }
    }
    strided_write<T, M, KEEP, BSYNC>(ch, role, type, base, s, safe, r0, xf, xl, gval, tin, out,
                                     tout, merges, stores, keep, kfs, krs);
}

// ---------------------------------------------------------------------------
// Strided lines (X or Y sweep).  A work-group owns LPG consecutive z lines and
// every row of them (MAXT = 32 * LPG threads for lines of up to 256 rows in
// chunks of M = 8); MINW is the occupancy the register allocation must allow
// (waves per SIMD), so that two half-width work-groups can share a CU.
// BSYNC: the state holds the node values on BOUND/VALVE nodes with no-slip
// BCs, so the Dirichlet values come from the Picard layer (else from the node
// values; that case, only met by a step without a boundary fill after the
// caller replaced the state or node values, has one instantiation per tile
// shape, without the LDS-parked layer); stores: bit 0 writes the sweep output, bit 1 the merged Picard
// layer (a sweep whose output or Picard layer nothing reads skips it), kStoreStream streams both
// past the caches.
template <class T, int AXIS, int M, int LPG, int MAXT, int MINW, bool KEEP, bool BSYNC>
__global__ void __launch_bounds__(MAXT, MINW) fused_strided_kernel(
    const InstDesc<T>* __restrict__ insts, T three_dt, const uint8_t* __restrict__ type,
    const T* __restrict__ gv0, const T* __restrict__ gv1, const T* __restrict__ gv2,
    const T* __restrict__ gv3, Layer4<T> in, Layer4<T> tin, Layer4<T> out, Layer4<T> tout,
    int merges, int stores, const int32_t* __restrict__ work, int kmax) {
    const T* gval[4] = {gv0, gv1, gv2, gv3};
    strided_tile<T, AXIS, M, LPG, KEEP, BSYNC>(insts, three_dt, type, gval, in, tin, out, tout,
                                               merges, stores, work, kmax);
}

// Contiguous lines (Z sweep): 32 lanes per line, two lines per wavefront.  A
// lane's chunk is contiguous in memory, so it is read and written with 16-byte
// vector accesses (the wave covers its two lines with back-to-back vectors);
// the along-line neighbours of a chunk come from the adjacent lanes by shuffle
// and the cross-line derivatives of w from vector loads of the neighbour lines.
// The reduced system of a line never leaves its lanes (no barrier).  PARK: the
// lane's Picard-layer values wait for the merge in LDS (16-byte slots,
// lane-contiguous) instead of registers.  LPL = 64: a line spans a whole
// wavefront (half the rows per lane, one line per wave, twice the waves per
// work-group) for fewer registers per lane.  VALL: every run of the launch is
// whole and 16-byte aligned, so a lane reads its rows with vector loads from
// a masked address whatever its row count (0 or M), and all of a lane's loads
// go out in one batch (a load used under a condition becomes a branch that
// waits for it).
template <class T, int M, int MINW, bool PARK, int LPL, bool VALL>
__global__ void __launch_bounds__(64 * kZWaves * (LPL / 32), MINW) fused_z_kernel(
    const InstDesc<T>* __restrict__ insts, T three_dt, const uint8_t* __restrict__ type,
    const T* __restrict__ gv0, const T* __restrict__ gv1, const T* __restrict__ gv2,
    const T* __restrict__ gv3, Layer4<T> in, Layer4<T> tin, Layer4<T> out, Layer4<T> tout,
    int merges, int bsync, int stores, const int32_t* __restrict__ lines,
    const int32_t* __restrict__ winst) {
    extern __shared__ char smem_raw[];
    const T* gv[4] = {gv0, gv1, gv2, gv3};
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    // every line of a work-group belongs to one instance
    const int b = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int inst = winst[b];
    const Geom g = insts[inst].g;
    SweepParams<T> p = insts[inst].p[AX_Z];
    p.three_dt = three_dt;

    // 2 * kZWaves slots per work-group, 64 / LPL per wavefront.  A slot holds a
    // run of `cnt` consecutive z lines starting at line x*ny + y (-1 = none):
    // adjacent in memory, they form one block-diagonal system (every line
    // starts and ends on a non-IN face node), solved as one line of cnt * nz
    // rows -- several short lines of a batch share the lanes sized for the
    // longest.
    const int sub = lane / LPL, ll = lane % LPL;
    const int slot = b * (2 * kZWaves) + (64 / LPL) * wave + sub;
    const int lid = lines[2 * slot], cnt = lines[2 * slot + 1];
    const bool live = lid >= 0;
    const int n = live ? cnt * g.nz : 0;
    const int used = (n + M - 1) / M;  // lanes of the slot holding rows (<= LPL)
    const int64_t base = live ? insts[inst].off + (int64_t)lid * g.nz : 0;
    const int r0 = ll * M;
    const bool ok = live && ll < used;
    const int valid = ok ? min(M, n - r0) : 0;
    const bool vec = VALL && valid == M;
    const int64_t c0 = base + r0;

    // inputs: node types (with the BC kinds), the node values when the state
    // is out of sync with them, the state (in), the Picard layer (tin) and the
    // neighbour lines of w (cross-line derivatives; the active lines have x, y
    // in [1, n-2], so those lines exist)
    int tyr[M];
    T gb[4][M], cu[4][M], tv[4][M], wa[M], wb[M], wc[M], wd[M];
    {
        const T* tw = tin.f[2];
        if constexpr (VALL) {
            load_bytes_raw<M, true>(type, c0, valid, tyr);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                load_run_raw<T, M, true>(in.f[v], c0, valid, cu[v]);
                load_run_raw<T, M, true>(tin.f[v], c0, valid, tv[v]);
            }
            load_run_raw<T, M, true>(tw + (ok ? g.sx : 0), c0, valid, wa);
            load_run_raw<T, M, true>(tw - (ok ? g.sx : 0), c0, valid, wb);
            load_run_raw<T, M, true>(tw + (ok ? g.sy : 0), c0, valid, wc);
            load_run_raw<T, M, true>(tw - (ok ? g.sy : 0), c0, valid, wd);
            if (!bsync) {
#pragma unroll
                for (int v = 0; v < 4; v++) load_run_raw<T, M, true>(gv[v], c0, valid, gb[v]);
            }
#pragma unroll
            for (int j = 0; j < M; j++) {
                pin(tyr[j]);
                pin(wa[j]);
                pin(wb[j]);
                pin(wc[j]);
                pin(wd[j]);
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    pin(cu[v][j]);
                    pin(tv[v][j]);
                }
            }
            if (!bsync) {
#pragma unroll
                for (int j = 0; j < M; j++)
#pragma unroll
                    for (int v = 0; v < 4; v++) pin(gb[v][j]);
            }
        } else {
#pragma unroll
            for (int j = 0; j < M; j++) tyr[j] = j < valid ? type[c0 + j] : kOutside;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                load_run<T, M>(in.f[v] + c0, cu[v], vec, valid);
                load_run<T, M>(tin.f[v] + c0, tv[v], vec, valid);
            }
            load_run<T, M>(tw + c0 + (ok ? g.sx : 0), wa, vec, valid);
            load_run<T, M>(tw + c0 - (ok ? g.sx : 0), wb, vec, valid);
            load_run<T, M>(tw + c0 + (ok ? g.sy : 0), wc, vec, valid);
            load_run<T, M>(tw + c0 - (ok ? g.sy : 0), wd, vec, valid);
        }
    }

    // node types of the chunk and of the rows either side (from the neighbour lanes)
    uint8_t ty[M + 2];
#pragma unroll
    for (int j = 0; j < M; j++) ty[j + 1] = j < valid ? (uint8_t)tyr[j] : kOutside;
    {
        int prev = lane_prev((int)ty[M]);
        int next = lane_next((int)ty[1]);
        ty[0] = (ll > 0 && ok) ? (uint8_t)prev : kOutside;
        ty[M + 1] = (ll < LPL - 1 && ok && r0 + M < n) ? (uint8_t)next : kOutside;
    }
    int role[M];
#pragma unroll
    for (int j = 0; j < M; j++) role[j] = row_role(ty[j + 1], ty[j], ty[j + 2]);
    // Dirichlet values of the boundary rows (ApplyBC0/1; the BC kinds are in
    // the flags bytes): the Picard layer while the state is in sync with the
    // node values (it holds them off NODE_IN), else the node values of the
    // whole chunk, one vector load per field (a load per boundary row would
    // sit inside the elimination chain)
    if constexpr (!VALL) {
#pragma unroll
        for (int j = 0; j < M; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) gb[v][j] = 0;
        bool anyb = false;
#pragma unroll
        for (int j = 0; j < M; j++) anyb = anyb || (role[j] != RL_IN && role[j] != RL_NONE);
        if (!bsync && anyb) {
#pragma unroll
            for (int v = 0; v < 4; v++) load_run<T, M>(gv[v] + c0, gb[v], vec, valid);
        }
    }
    if (bsync) {
#pragma unroll
        for (int j = 0; j < M; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) gb[v][j] = 0;
    }
    T wx[M], wy[M];
#pragma unroll
    for (int j = 0; j < M; j++) {
        wx[j] = wa[j] - wb[j];
        wy[j] = wc[j] - wd[j];
    }
    T tprev[4], tnext[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        tprev[v] = lane_prev(tv[v][M - 1]);
        tnext[v] = lane_next(tv[v][0]);
    }
    using V16 = typename Vec16<T>::type;
    constexpr int VW = Vec16<T>::n;
    constexpr int NS = 4 * ((M + VW - 1) / VW);  // 16-byte slots per lane
    V16* park = reinterpret_cast<V16*>(smem_raw) + (int64_t)wave * NS * 64 + lane;
    if (PARK) {
#pragma unroll
        for (int v = 0; v < 4; v++)
#pragma unroll
            for (int k = 0; k < (M + VW - 1) / VW; k++) {
                V16 x;
                T* e = reinterpret_cast<T*>(&x);
#pragma unroll
                for (int q = 0; q < VW; q++) e[q] = k * VW + q < M ? tv[v][k * VW + q] : (T)0;
                park[(v * ((M + VW - 1) / VW) + k) * 64] = x;
            }
    }

    Chunk<T, M> ch;
    Row<T> rfirst, rlast;
#pragma unroll
    for (int j = 0; j < M; j++) {
        // correction form (see interior_row_delta): Picard layer around the row
        T wm[4], w0[4], wp[4], ci[4], gd[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            wm[v] = j > 0 ? tv[v][j - 1] : tprev[v];
            w0[v] = tv[v][j];
            wp[v] = j < M - 1 ? tv[v][j + 1] : tnext[v];
            ci[v] = cu[v][j];
            gd[v] = bsync ? w0[v] : gb[v][j];
        }
        Row<T> ri, rb, r;
        interior_row_delta<T, AX_Z>(p, wm, w0, wp, ci, wx[j], wy[j], ri);
        boundary_row_delta<T>(role[j], ty[j + 1], wm, w0, wp, gd, rb);
        const bool isin = role[j] == RL_IN;
        r.av = isin ? ri.av : rb.av; r.bv = isin ? ri.bv : rb.bv; r.cv = isin ? ri.cv : rb.cv;
        r.at = isin ? ri.at : rb.at; r.bt = isin ? ri.bt : rb.bt; r.ct = isin ? ri.ct : rb.ct;
#pragma unroll
        for (int v = 0; v < 4; v++) r.d[v] = isin ? ri.d[v] : rb.d[v];
        if (j == 0) rfirst = r;
        else if (j == M - 1) rlast = r;
        else chunk_down<T, M>(ch, j, r);
    }
    if (M >= 3) chunk_up<T, M>(ch);
    // reduced rows (identity for lanes without rows: their rows are all outside)
    // and the line's reduced solve inside its 32 lanes
    T xf[4], xl[4];
    {
        T f[kRedFields], l[kRedFields];
        chunk_reduced<T, M>(ch, rfirst, rlast, f, l);
        solve_chunks_in_lanes<T>(f, l, ll, used, LPL, xf, xl);
    }
    if (!ok) return;
    if (PARK) {
        // the parked copy is read back (not the registers of the loads)
        asm volatile("" ::: "memory");
#pragma unroll
        for (int v = 0; v < 4; v++)
#pragma unroll
            for (int k = 0; k < (M + VW - 1) / VW; k++) {
                const V16 x = park[(v * ((M + VW - 1) / VW) + k) * 64];
                const T* e = reinterpret_cast<const T*>(&x);
#pragma unroll
                for (int q = 0; q < VW; q++)
                    if (k * VW + q < M) tv[v][k * VW + q] = e[q];
            }
    }
    const T km = merges > 1 ? (T)0.75 : (T)0.5;
    with_store_hint((stores & kStoreStream) != 0, [&](auto ntc) {
        constexpr bool NT = decltype(ntc)::value;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            T x[M], t[M];
#pragma unroll
            for (int j = 0; j < M; j++) {
                T d;  // correction of row j (xf / xl: of the chunk's first / last row)
                if (j == 0) d = xf[v];
                else if (j == M - 1) d = xl[v];
                else {
                    const T F = v < 3 ? ch.Fv[j] : ch.Ft[j];
                    const T G = v < 3 ? ch.Gv[j] : ch.Gt[j];
                    d = ch.E[v][j] + F * xf[v] + G * xl[v];
                }
                x[j] = tv[v][j] + d;
                if (!bsync && dirichlet_row(role[j], ty[j + 1], v < 3 ? 0 : 1)) x[j] = gb[v][j];
                // rows no segment covers get the sweep input's value
                if (role[j] == RL_NONE) x[j] = cu[v][j];
                t[j] = role[j] == RL_IN ? tv[v][j] + km * d : tv[v][j];
            }
            // whole runs: partial-vector stores would cost far more than the bytes
            if (stores & 1) store_run<T, M, NT>(out.f[v] + c0, x, vec, valid);
            if (stores & 2) store_run<T, M, NT>(tout.f[v] + c0, t, vec, valid);
        }
    });
}

// Strided-sweep variants: tile width (bytes of z lines per work-group: 128 =
// one work-group per CU, 64 = two, 32 = four, in different phases of load,
// elimination and reduced solve), rows per chunk (8, or 4 for lines of at most
// 128 nodes: twice the lanes per line; the reduced system keeps at most 64
// rows), and whether the Picard layer read for the coefficients is parked in
// LDS for the merge instead of re-read from memory (variants 6-9: 0, 1, 2 and
// 3 with the parked layer).  Registers are capped at 4 (FP32) / 2 (FP64)
// waves per SIMD.
struct StridedShape {
    int bytes, chunk;
    bool keep;
};
constexpr StridedShape kShapes[kFusedVariants] = {
    {128, 8, false}, {64, 8, false}, {32, 8, false}, {128, 4, false}, {64, 4, false},
    {32, 4, false},  {128, 8, true}, {64, 8, true},  {32, 8, true},   {128, 4, true}};
constexpr int kChunk = 8;
template <class T>
constexpr int variant_chunk(int variant) { return kShapes[variant].chunk; }
template <class T>
constexpr int variant_lpg(int variant) { return kShapes[variant].bytes / (int)sizeof(T); }

int strided_waves_for(int n, int lpg, int m = kChunk) {
    const int gpw = 64 / lpg;
    return (n + m * gpw - 1) / (m * gpw);
}

// Rows per lane of a Z line spread over `lpl` lanes (0: too long).  32-lane
// lines take up to 16 rows per lane, 64-lane lines up to 8.
int z_chunk(int n, int lpl = 32) {
    for (int m : {2, 4, 8, 16})
        if (lpl * m >= n && (lpl == 32 || m <= 8)) return m;
    return 0;
}

}  // namespace

template <class T>
int strided_lpg(int variant) {
    return variant_lpg<T>(variant);
}

template <class T>
int strided_chunk_rows(int variant) {
    return variant_chunk<T>(variant);
}

constexpr size_t kMaxLds = 160 * 1024;

template <class T>
size_t strided_lds(int n, int v) {
    const int lpg = variant_lpg<T>(v), m = variant_chunk<T>(v);
    const int P = strided_waves_for(n, lpg, m) * (64 / lpg);
    if (!kShapes[v].keep) return (size_t)lpg * (kRedFields * 2 * P + 1) * sizeof(T);
    return (size_t)lpg * kNormFields * pow2_at_least(P) * sizeof(T) +
           (size_t)4 * P * m * keep_row_pitch<T>(lpg, m) * sizeof(T);
}

// A strided variant fits lines of n nodes when its work-group (one lane per
// chunk of every line) stays within 32 * LPG threads and its LDS within a CU.
template <class T>
bool strided_variant_fits(int n, int v) {
    const int lpg = variant_lpg<T>(v);
    return strided_waves_for(n, lpg, variant_chunk<T>(v)) * 64 <= 32 * lpg &&
           strided_lds<T>(n, v) <= kMaxLds;
}

template <class T>
bool fused_supported(const Geom& g) {
    if (z_chunk(g.nz) == 0 || (sizeof(T) == 8 && z_chunk(g.nz) > 8)) return false;
    // the chunk-8, memory-read variants must fit (the others are optional)
    for (int v = 0; v < 3; v++)
        if (!strided_variant_fits<T>(g.nx, v) || !strided_variant_fits<T>(g.ny, v)) return false;
    return true;
}

template <class T>
bool fused_variant_ok(int n, int v) {
    return strided_variant_fits<T>(n, v);
}

template <class T>
void launch_fused_sweep(const InstDesc<T>* insts, T three_dt, int axis, const FusedWork& w,
                        const uint8_t* tflags, const T* const* gval, Layer4<T> in, Layer4<T> tin,
                        Layer4<T> out, Layer4<T> tout, int merges, int bsync, int stores,
                        int variant, hipStream_t st) {
    if (w.nblocks == 0) return;
    const unsigned blocks = (unsigned)w.nblocks;
    if (w.stream) stores |= kStoreStream;
    if (axis == AX_Z) {
        // variants 1, 3, 5, 7 cap the registers for 4 (FP32) / 2 (FP64) waves per
        // SIMD; 2, 3, 6, 7 park the Picard layer in LDS; 4-7 spread a line over
        // a whole wavefront
        constexpr int MW = sizeof(T) == 8 ? 2 : 4;
        const bool park = (variant & 2) != 0, wide = (variant & 4) != 0;
        const int lpl = wide ? 64 : 32;
        const int m = z_chunk(w.maxn, lpl);
        if (m == 0) throw std::string("fused sweep: Z variant does not fit the line length");
        const int vw = 16 / (int)sizeof(T);
        const int nthreads = 64 * kZWaves * (lpl / 32);
        const size_t lds = park ? (size_t)(nthreads / 64) * 64 * 4 * ((m + vw - 1) / vw) * 16 : 0;
        // whole aligned runs: every line 16-byte aligned and a multiple of m rows
        const bool vall = w.vec_ok && w.nz_gcd % m == 0;
#define LZV(MM, W, PK, LL, VA)                                                                       \
    hipLaunchKernelGGL((fused_z_kernel<T, MM, W, PK, LL, VA>), dim3(blocks), dim3(nthreads), lds, \
                       st, insts, three_dt, tflags, gval[0], gval[1], gval[2], gval[3], in, tin,   \
                       out, tout, merges, bsync, stores, w.d_items, w.d_inst)
#define LZ(MM, W, PK, LL)              \
    if (vall) LZV(MM, W, PK, LL, true); \
    else LZV(MM, W, PK, LL, false)
#define LZM(W, PK)                      \
    if (!wide) {                        \
        if (m == 2) LZ(2, W, PK, 32);   \
        else if (m == 4) LZ(4, W, PK, 32); \
        else if (m == 8) LZ(8, W, PK, 32); \
        else LZ(16, 1, PK, 32);         \
    } else {                            \
        if (m == 2) LZ(2, W, PK, 64);   \
        else if (m == 4) LZ(4, W, PK, 64); \
        else LZ(8, W, PK, 64);          \
    }
        if ((variant & 3) == 0) {
            LZM(1, false)
        } else if ((variant & 3) == 1) {
            LZM(MW, false)
        } else if ((variant & 3) == 2) {
            LZM(1, true)
        } else {
            LZM(MW, true)
        }
#undef LZM
#undef LZ
#undef LZV
    } else {
        const int lpg = variant_lpg<T>(variant);
        const int mch = variant_chunk<T>(variant);
        if (!strided_variant_fits<T>(w.maxn, variant))
            throw std::string("fused sweep: tile variant does not fit the line length");
        const int waves = strided_waves_for(w.maxn, lpg, mch);
        size_t lds = strided_lds<T>(w.maxn, variant);
#define LS(AX, V, BS)                                                                             \
    hipLaunchKernelGGL((fused_strided_kernel<T, AX, kShapes[V].chunk, kShapes[V].bytes / (int)sizeof(T), \
                                             32 * (kShapes[V].bytes / (int)sizeof(T)),            \
                                             16 / (int)sizeof(T), kShapes[V].keep, BS>),          \
                       dim3(blocks), dim3(64 * waves), lds, st, insts, three_dt, tflags, gval[0], \
                       gval[1], gval[2], gval[3], in, tin, out, tout, merges, stores, w.d_items, \
                       w.kmax)
#define LSA(AX)                                 \
    if (bsync) {                                \
        switch (variant) {                      \
            case 0: LS(AX, 0, true); break;     \
            case 1: LS(AX, 1, true); break;     \
            case 2: LS(AX, 2, true); break;     \
            case 3: LS(AX, 3, true); break;     \
            case 4: LS(AX, 4, true); break;     \
            case 5: LS(AX, 5, true); break;     \
            case 6: LS(AX, 6, true); break;     \
            case 7: LS(AX, 7, true); break;     \
            case 8: LS(AX, 8, true); break;     \
            default: LS(AX, 9, true); break;    \
        }                                       \
    } else {                                    \
        switch (oos) {                          \
            case 0: LS(AX, 0, false); break;    \
            case 1: LS(AX, 1, false); break;    \
            case 2: LS(AX, 2, false); break;    \
            case 3: LS(AX, 3, false); break;    \
            case 4: LS(AX, 4, false); break;    \
            default: LS(AX, 5, false); break;   \
        }                                       \
    }
        // out of sync: the variant of the same tile shape without the parked
        // layer (the same work list: it depends on the tile width only)
        const int oos = kShapes[variant].keep ? variant - 6 : variant;
        if (!bsync) lds = strided_lds<T>(w.maxn, oos);
        if (axis == AX_X) {
            LSA(AX_X);
        } else {
            LSA(AX_Y);
        }
#undef LSA
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

// Work of one axis over all instances.  The work-group of a launch is sized
// for the longest line of the batch (`maxn`); grids with shorter lines pack
// several tiles (X/Y) or line runs (Z) into one work-group, so that short
// lines do not run in lanes sized for long ones.
//   Z: slots of (first line x*ny + y, count): runs of consecutive lines
//      holding a NODE_IN node, at most floor(capacity / nz) lines each
//      (capacity = 32 lanes x the rows per lane of the longest line), each
//      instance's slots padded with (-1, 0) to whole work-groups (2 per wave),
//      plus the instance of every work-group.
//   X/Y: items of kmax + 1 ints: (instance, tile, tile, ...), tile = outer *
//      ztiles + z tile of the tiles of lpg z lines holding a NODE_IN node,
//      -1-padded; an instance takes floor(P / chunks per line) tiles per item
//      (P = chunks per z-line slot of the work-group, chunks of m rows).
void build_fused_work(const uint8_t* h_type, const std::vector<Geom>& geoms,
                      const std::vector<int64_t>& offs, int axis, int lpg, int m, size_t elem,
                      FusedWork& w) {
    w.release();
    std::vector<int32_t> items, winst;
    {
        // stream the stores of sweeps whose fields exceed the memory-side
        // cache (16 fields move per sweep)
        int64_t cells = 0;
        for (const Geom& g : geoms) cells += (int64_t)g.nx * g.ny * g.nz;
        w.stream = (double)cells * (double)elem * 16.0 > (double)kStreamStoreBytes ? 1 : 0;
    }
    w.maxn = 0;
    w.vec_ok = 1;
    w.nz_gcd = 0;
    w.kmax = 1;
    for (const Geom& g : geoms)
        w.maxn = std::max(w.maxn, axis == AX_Z ? g.nz : axis == AX_X ? g.nx : g.ny);
    if (axis == AX_Z) {
        const int cap = 32 * std::max(z_chunk(w.maxn, 32), 2);
        for (size_t i = 0; i < geoms.size(); i++) {
            const Geom& g = geoms[i];
            const uint8_t* t = h_type + offs[i];
            if (((size_t)g.nz * elem) % 16 || ((size_t)offs[i] * elem) % 16) w.vec_ok = 0;
            w.nz_gcd = std::gcd(w.nz_gcd, g.nz);
            const int K = std::max(1, cap / g.nz);
            const size_t first = items.size();
            int run0 = -1, cnt = 0;
            for (int x = 0; x < g.nx; x++)
                for (int y = 0; y < g.ny; y++) {
                    const uint8_t* row = &t[x * g.sx + y * g.sy];
                    bool any = false;
                    for (int z = 0; z < g.nz && !any; z++) any = row[z] == NODE_IN;
                    if (!any) continue;
                    const int lid = x * g.ny + y;
                    if (cnt > 0 && lid == run0 + cnt && cnt < K) {
                        cnt++;
                        continue;
                    }
                    if (cnt > 0) {
                        items.push_back(run0);
                        items.push_back(cnt);
                    }
                    run0 = lid;
                    cnt = 1;
                }
            if (cnt > 0) {
                items.push_back(run0);
                items.push_back(cnt);
            }
            const size_t per = 2 * 2 * kZWaves;
            while ((items.size() - first) % per) {
                items.push_back(-1);
                items.push_back(0);
            }
            for (size_t k = first; k < items.size(); k += per) winst.push_back((int32_t)i);
        }
        w.nblocks = (int64_t)winst.size();
    } else {
        const int P = strided_waves_for(w.maxn, lpg, m) * (64 / lpg);
        std::vector<std::vector<int32_t>> tiles(geoms.size());
        std::vector<int> per(geoms.size());
        for (size_t i = 0; i < geoms.size(); i++) {
            const Geom& g = geoms[i];
            const uint8_t* t = h_type + offs[i];
            const int n = axis == AX_X ? g.nx : g.ny;
            per[i] = std::max(1, P / ((n + m - 1) / m));
            w.kmax = std::max(w.kmax, per[i]);
            const int ztiles = (g.nz + lpg - 1) / lpg;
            const int nouter = axis == AX_X ? g.ny : g.nx;
            const int64_t so = axis == AX_X ? g.sy : g.sx, sr = axis == AX_X ? g.sx : g.sy;
            for (int o = 0; o < nouter; o++)
                for (int zt = 0; zt < ztiles; zt++) {
                    bool any = false;
                    for (int r = 0; r < n && !any; r++)
                        for (int z = zt * lpg; z < std::min(g.nz, (zt + 1) * lpg) && !any; z++)
                            any = t[o * so + r * sr + z] == NODE_IN;
                    if (any) tiles[i].push_back(o * ztiles + zt);
                }
        }
        const int W = w.kmax + 1;
        for (size_t i = 0; i < geoms.size(); i++)
            for (size_t k = 0; k < tiles[i].size(); k += per[i]) {
                items.push_back((int32_t)i);
                for (int q = 0; q < w.kmax; q++)
                    items.push_back(q < per[i] && k + q < tiles[i].size() ? tiles[i][k + q] : -1);
            }
        w.nblocks = (int64_t)items.size() / W;
        if (axis == AX_Y) {
            // The Y sweep reads what the Z sweep wrote: x-planes, in list order
            // per XCD.  Walking its own list backwards, it starts on the planes Z
            // wrote last, which the memory-side cache still holds.
            const int64_t nb = w.nblocks;
            for (int64_t k = 0; k < nb / 2; k++)
                for (int q = 0; q < W; q++) std::swap(items[k * W + q], items[(nb - 1 - k) * W + q]);
        }
    }
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

void build_node_flags(const uint8_t* t, const uint8_t* bv, const uint8_t* bt, int64_t n,
                      uint8_t* out) {
    for (int64_t i = 0; i < n; i++)
        out[i] = t[i] == NODE_IN ? (uint8_t)0
                                 : (uint8_t)(t[i] | (bv[i] == BC_FREE ? kFlagFreeVel : 0) |
                                             (bt[i] == BC_FREE ? kFlagFreeT : 0));
}

template bool fused_supported<float>(const Geom&);
template bool fused_supported<double>(const Geom&);
template bool fused_variant_ok<float>(int, int);
template bool fused_variant_ok<double>(int, int);
template int strided_lpg<float>(int);
template int strided_lpg<double>(int);
template int strided_chunk_rows<float>(int);
template int strided_chunk_rows<double>(int);
template void launch_fused_sweep<float>(const InstDesc<float>*, float, int, const FusedWork&,
                                        const uint8_t*, const float* const*, Layer4<float>,
                                        Layer4<float>, Layer4<float>, Layer4<float>, int, int, int,
                                        int, hipStream_t);
template void launch_fused_sweep<double>(const InstDesc<double>*, double, int, const FusedWork&,
                                         const uint8_t*, const double* const*, Layer4<double>,
                                         Layer4<double>, Layer4<double>, Layer4<double>, int, int,
                                         int, int, hipStream_t);

}  // namespace cmcx
