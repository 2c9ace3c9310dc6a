// Plain segmented Thomas sweep: one thread per (segment, field), one launch per axis.
//
// This is the per-axis-launch baseline.  Each thread assembles its segment's
// rows on the fly (BuildMatrix, src/FluidSolver3D/AdiSolver3D.cpp:732-802; BC
// rows ApplyBC0/1, :804-852), runs the Thomas recurrence of
// src/Common/Algorithms.h:21-38 with c' and d' kept in a global scratch buffer
// laid out (row, field, segment) so that consecutive threads touch consecutive
// words, then back-substitutes, writing every node of the segment into `out`
// (UpdateSegment, :707-730) and the Picard merge temp_out = (temp_in + x)/2 on
// the segment's NODE_IN nodes (TimeLayer3D.h:415-436).  A node that ends one
// segment and starts the next on the same line (a one-node wall) is written by
// the later segment only, so the result does not depend on thread timing (the
// reference's threads race there when the wall is BC_FREE).
//
// Arithmetic follows the reference's operation order and is compiled without
// contraction; with IEEE division it reproduces the reference's CPU arithmetic.
#pragma clang fp contract(off)

#include "cmcx_internal.h"

namespace cmcx {

namespace {

// RHS of the row at node c for field `var` of a sweep along `axis`
// (AdiSolver3D.cpp:755-801; stencils TimeLayer3D.h:338-343, 554-588).
template <class T>
__device__ __forceinline__ T row_rhs(int axis, int var, const T* __restrict__ curv, int64_t c,
                                     const T* __restrict__ tu, const T* __restrict__ tv,
                                     const T* __restrict__ tw, const T* __restrict__ tt,
                                     const Geom& g, const SweepParams<T>& p) {
    T base = curv[c] * 3 / p.dt;
    const int64_t sx = g.sx, sy = g.sy;
    if (var == 3) {
        T phi;
        if (axis == AX_X) {
            T u_x = (tu[c + sx] - tu[c - sx]) / (2 * p.hx);
            T v_x = (tv[c + sx] - tv[c - sx]) / (2 * p.hx);
            T w_x = (tw[c + sx] - tw[c - sx]) / (2 * p.hx);
            T u_y = (tu[c + sy] - tu[c - sy]) / (2 * p.hy);
            T u_z = (tu[c + 1] - tu[c - 1]) / (2 * p.hz);
            phi = 2 * u_x * u_x + v_x * v_x + w_x * w_x + v_x * u_y + w_x * u_z;
        } else if (axis == AX_Y) {
            T u_y = (tu[c + sy] - tu[c - sy]) / (2 * p.hy);
            T v_y = (tv[c + sy] - tv[c - sy]) / (2 * p.hy);
            T w_y = (tw[c + sy] - tw[c - sy]) / (2 * p.hy);
            T v_x = (tv[c + sx] - tv[c - sx]) / (2 * p.hx);
            T v_z = (tv[c + 1] - tv[c - 1]) / (2 * p.hz);
            phi = u_y * u_y + 2 * v_y * v_y + w_y * w_y + u_y * v_x + w_y * v_z;
        } else {
            T u_z = (tu[c + 1] - tu[c - 1]) / (2 * p.hz);
            T v_z = (tv[c + 1] - tv[c - 1]) / (2 * p.hz);
            T w_z = (tw[c + 1] - tw[c - 1]) / (2 * p.hz);
            T w_x = (tw[c + sx] - tw[c - sx]) / (2 * p.hx);
            T w_y = (tw[c + sy] - tw[c - sy]) / (2 * p.hy);
            phi = u_z * u_z + v_z * v_z + 2 * w_z * w_z + u_z * w_x + v_z * w_y;
        }
        return base + p.t_phi * phi;
    }
    if (var == axis) {
        int64_t s = axis == AX_X ? sx : axis == AX_Y ? sy : 1;
        T grad = (tt[c + s] - tt[c - s]) / (2 * p.h);
        return base - p.v_T * grad;
    }
    return base;
}

template <class T>
__global__ void __launch_bounds__(256) plain_sweep_kernel(
    Geom g, int axis, const int32_t* __restrict__ segs, int64_t nseg, SweepParams<T> p,
    const uint8_t* __restrict__ type, const uint8_t* __restrict__ bcv, const uint8_t* __restrict__ bct, const T* __restrict__ gv0,
    const T* __restrict__ gv1, const T* __restrict__ gv2, const T* __restrict__ gt,
    Layer4<T> in, Layer4<T> tin, Layer4<T> out, Layer4<T> tout, int merges, int var0,
    T* __restrict__ sc, T* __restrict__ sd) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    const int var = var0 + (int)blockIdx.y;
    const int px = segs[s], py = segs[nseg + s], pz = segs[2 * nseg + s];
    const int ex = segs[3 * nseg + s], ey = segs[4 * nseg + s], ez = segs[5 * nseg + s];
    const int n = segs[6 * nseg + s];
    const int64_t stride = axis == AX_X ? g.sx : axis == AX_Y ? g.sy : 1;
    const int64_t c0 = px * g.sx + py * g.sy + pz;
    const int64_t c1 = ex * g.sx + ey * g.sy + ez;
    const T* __restrict__ gval = var == 0 ? gv0 : var == 1 ? gv1 : var == 2 ? gv2 : gt;
    const uint8_t* __restrict__ bcs = var == 3 ? bct : bcv;
    const T vis = var == 3 ? p.vis_t : p.vis_v;
    const T* adv = tin.f[axis];  // may alias tout (in-place merge order)
    const T* __restrict__ curv = in.f[var];
    const int64_t ld = 4 * nseg;  // scratch pitch of one row
    T* __restrict__ cp = sc + (int64_t)var * nseg + s;
    T* __restrict__ dp = sd + (int64_t)var * nseg + s;

    // row 0: ApplyBC0 at the first node, then c0/b0, d0/b0
    T b, c, d, a;
    if (bcs[c0] == BC_FREE) { b = 2.0; c = -1.0; d = 0.0; }
    else { b = 1.0; c = 0.0; d = gval[c0]; }
    T cprev = c / b;
    T dprev = d / b;
    cp[0] = cprev;
    dp[0] = dprev;
    const T bb = 3 / p.dt + 2 * vis;
    const T two_h = 2 * p.h;
    int64_t cell = c0;
    for (int r = 1; r < n - 1; r++) {
        cell += stride;
        T wv = adv[cell];
        a = -wv / two_h - vis;
        c = wv / two_h - vis;
        d = row_rhs<T>(axis, var, curv, cell, tin.f[0], tin.f[1], tin.f[2], tin.f[3], g, p);
        T den = bb - a * cprev;
        cprev = c / den;
        dprev = (d - dprev * a) / den;
        cp[r * ld] = cprev;
        dp[r * ld] = dprev;
    }
    // row n-1: ApplyBC1 at the last node (c[n-1] = 0)
    if (bcs[c1] == BC_FREE) { a = -1.0; b = 2.0; d = 0.0; }
    else { a = 0.0; b = 1.0; d = gval[c1]; }
    T x = (d - dprev * a) / (b - a * cprev);
    T* __restrict__ o = out.f[var];
    const T* ti = tin.f[var];
    T* to = tout.f[var];
    const int epos = axis == AX_X ? ex : axis == AX_Y ? ey : ez;
    const int dim = axis == AX_X ? g.nx : axis == AX_Y ? g.ny : g.nz;
    const bool shared_end = epos + 1 < dim && type[c1 + stride] == NODE_IN;
    if (!shared_end) o[c1] = x;
    cell = c1;
    for (int r = n - 2; r >= 1; r--) {
        cell -= stride;
        x = dp[r * ld] - cp[r * ld] * x;
        o[cell] = x;
        T m = (ti[cell] + x) / 2;
        if (merges > 1) m = (m + x) / 2;
        to[cell] = m;
    }
    x = dp[0] - cp[0] * x;
    o[c0] = x;
}

}  // namespace

template <class T>
void launch_plain_sweep(const Geom& g, int axis, const SegList& segs, const SweepParams<T>& p,
                        const uint8_t* type, const uint8_t* bcv, const uint8_t* bct,
                        const T* const* gval, Layer4<T> in, Layer4<T> tin, Layer4<T> out,
                        Layer4<T> tout, int merges, int var0, int var1, T* scratch,
                        hipStream_t st) {
    if (segs.n == 0) return;
    const int bs = 256;
    dim3 grid((unsigned)((segs.n + bs - 1) / bs), (unsigned)(var1 - var0));
    // scratch: c' block then d' block, each max_size * 4 * nseg
    int64_t half = (int64_t)segs.max_size * 4 * segs.n;
    hipLaunchKernelGGL((plain_sweep_kernel<T>), grid, dim3(bs), 0, st, g, axis, segs.d, segs.n, p,
                       type, bcv, bct, gval[0], gval[1], gval[2], gval[3], in, tin, out, tout, merges,
                       var0, scratch, scratch + half);
    CMCX_LAUNCH_CHECK();
}

template void launch_plain_sweep<float>(const Geom&, int, const SegList&, const SweepParams<float>&,
                                        const uint8_t*, const uint8_t*, const uint8_t*,
                                        const float* const*,
                                        Layer4<float>, Layer4<float>, Layer4<float>, Layer4<float>,
                                        int, int, int, float*, hipStream_t);
template void launch_plain_sweep<double>(const Geom&, int, const SegList&,
                                         const SweepParams<double>&, const uint8_t*,
                                         const uint8_t*, const uint8_t*, const double* const*,
                                         Layer4<double>,
                                         Layer4<double>, Layer4<double>, Layer4<double>, int, int,
                                         int, double*, hipStream_t);

}  // namespace cmcx
