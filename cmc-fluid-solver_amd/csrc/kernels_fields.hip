// Whole-grid field kernels of the time step: boundary fill, step set-up
// copies, divergence error, masked copy/clear and the mask checks.
#pragma clang fp contract(off)

#include "cmcx_internal.h"

#include <algorithm>

namespace cmcx {

namespace {

constexpr int kBS = 256;

inline unsigned blocks_for(int64_t n, int bs = kBS) { return (unsigned)((n + bs - 1) / bs); }

// UpdateBoundaries, CPU semantics (AdiSolver3D.cpp:290-294 -> TimeLayer3D.h:926-951):
// node velocity and temperature into the state on NODE_BOUND and NODE_VALVE nodes.
template <class T>
__global__ void update_boundaries_kernel(int64_t N, const uint8_t* __restrict__ type,
                                         const T* __restrict__ g0, const T* __restrict__ g1,
                                         const T* __restrict__ g2, const T* __restrict__ g3,
                                         Layer4<T> cur) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    uint8_t t = type[i];
    if (t == NODE_BOUND || t == NODE_VALVE) {
        cur.f[0][i] = g0[i];
        cur.f[1][i] = g1[i];
        cur.f[2][i] = g2[i];
        cur.f[3][i] = g3[i];
    }
}

// Step set-up (AdiSolver3D.cpp:310-311, 320): next <- cur on BOUND/VALVE nodes and
// temp <- cur.  The merged temp layer only changes on NODE_IN nodes, so both temp
// buffers get the non-IN nodes here; the NODE_IN nodes of the first sweep's temp are
// read straight from cur.  full != 0 copies every node into tA (in-place merge order).
template <class T>
__global__ void prep_step_kernel(int64_t N, const uint8_t* __restrict__ type, Layer4<T> cur,
                                 Layer4<T> next, Layer4<T> tA, Layer4<T> tB, int full) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    uint8_t t = type[i];
    if (full) {
#pragma unroll
        for (int f = 0; f < 4; f++) tA.f[f][i] = cur.f[f][i];
    } else if (t != NODE_IN) {
#pragma unroll
        for (int f = 0; f < 4; f++) {
            T v = cur.f[f][i];
            tA.f[f][i] = v;
            tB.f[f][i] = v;
        }
    }
    if (t == NODE_BOUND || t == NODE_VALVE) {
#pragma unroll
        for (int f = 0; f < 4; f++) next.f[f][i] = cur.f[f][i];
    }
}

// List forms of the two kernels above, over precomputed node lists (the mask is
// fixed): non-IN nodes for the step set-up, BOUND/VALVE nodes for the boundary
// fill.  They touch only the listed nodes instead of streaming the whole grid.
template <class T>
__global__ void update_boundaries_list_kernel(int64_t n, const uint32_t* __restrict__ ids,
                                              const T* __restrict__ g0, const T* __restrict__ g1,
                                              const T* __restrict__ g2, const T* __restrict__ g3,
                                              Layer4<T> cur) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = ids[t];
    cur.f[0][i] = g0[i];
    cur.f[1][i] = g1[i];
    cur.f[2][i] = g2[i];
    cur.f[3][i] = g3[i];
}

template <class T>
__global__ void prep_step_list_kernel(int64_t n, const uint32_t* __restrict__ ids,
                                      const uint8_t* __restrict__ type, Layer4<T> cur,
                                      Layer4<T> next, Layer4<T> tA, Layer4<T> tB) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = ids[t];
    const uint8_t ty = type[i];
    const bool bv = ty == NODE_BOUND || ty == NODE_VALVE;
#pragma unroll
    for (int f = 0; f < 4; f++) {
        T v = cur.f[f][i];
        tA.f[f][i] = v;
        tB.f[f][i] = v;
        if (bv) next.f[f][i] = v;
    }
}

template <class T>
__global__ void copy4_kernel(int64_t n, Layer4<T> dst, Layer4<T> src) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int f = 0; f < 4; f++) dst.f[f][i] = src.f[f][i];
}

// EvalDivError (TimeLayer3D.h:595-641): per NODE_IN node with i < nx-1, j < ny-1,
// k < nz-1, |flux divergence| of the dual cell; field-type sums, double result.
// Grid-stride loop, one (sum, count) partial per block, summed in block order by
// div_finish_kernel: the result does not depend on scheduling.
constexpr int kDivBlocks = 2048;

template <class T>
__global__ void __launch_bounds__(256) div_error_kernel(Geom g, const uint8_t* __restrict__ type,
                                                        const T* __restrict__ U,
                                                        const T* __restrict__ V,
                                                        const T* __restrict__ W, T hx, T hy, T hz,
                                                        double* __restrict__ part) {
    double val = 0.0, cnt = 0.0;
    const int64_t sx = g.sx, sy = g.sy;
    // a grid holds fewer than 2^31 nodes (checked at creation): 32-bit index math
    const uint32_t N = (uint32_t)g.N, nz = (uint32_t)g.nz, ny = (uint32_t)g.ny;
    for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < N; id += gridDim.x * blockDim.x) {
        const uint32_t line = id / nz;
        const int k = (int)(id - line * nz);
        const int i = (int)(line / ny);
        const int j = (int)(line - (uint32_t)i * ny);
        if (i < g.nx - 1 && j < g.ny - 1 && k < g.nz - 1 && type[id] == NODE_IN) {
            const int64_t c = id;
            T su = U[c] + U[c - sy] + U[c - sy - 1] + U[c - 1] - U[c - sx] - U[c - sx - sy] -
                   U[c - sx - sy - 1] - U[c - sx - 1];
            T sv = V[c] + V[c - sx] + V[c - sx - 1] + V[c - 1] - V[c - sy] - V[c - sx - sy] -
                   V[c - sx - sy - 1] - V[c - sy - 1];
            T sw = W[c] + W[c - sy] + W[c - sx - sy] + W[c - sx] - W[c - 1] - W[c - sy - 1] -
                   W[c - sx - sy - 1] - W[c - sx - 1];
            double ex = (double)(su * hz * hy) / 4.0;
            double ey = (double)(sv * hx * hz) / 4.0;
            double ez = (double)(sw * hx * hy) / 4.0;
            val += fabs(ex + ey + ez);
            cnt += 1.0;
        }
    }
    __shared__ double sv_[256], sc_[256];
    sv_[threadIdx.x] = val;
    sc_[threadIdx.x] = cnt;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sv_[threadIdx.x] += sv_[threadIdx.x + o];
            sc_[threadIdx.x] += sc_[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sv_[0];
        part[2 * blockIdx.x + 1] = sc_[0];
    }
}

// One wavefront: lane t sums the partials t, t + 64, ... in order, then a
// fixed butterfly over the lanes (the same order on every run).
__global__ void __launch_bounds__(64) div_finish_kernel(const double* __restrict__ part, int nb,
                                                        double* __restrict__ out) {
    const int t = threadIdx.x;
    double s = 0.0, c = 0.0;
    for (int b = t; b < nb; b += 64) {
        s += part[2 * b];
        c += part[2 * b + 1];
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        c += __shfl_xor(c, o, 64);
    }
    if (t == 0) {
        out[0] = s;
        out[1] = c;
    }
}

template <class T>
__global__ void scale_kernel(const T* __restrict__ src, T* __restrict__ dst, T a, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = a * src[i];
}

template <class T>
__global__ void convert_kernel(const double* __restrict__ src, T* __restrict__ dst, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (T)src[i];
}

template <class T>
__global__ void copy_typed_kernel(int64_t N, const uint8_t* __restrict__ type, int target,
                                  Layer4<T> src, Layer4<T> dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || type[i] != target) return;
#pragma unroll
    for (int f = 0; f < 4; f++) dst.f[f][i] = src.f[f][i];
}

template <class T>
__global__ void fill_typed_kernel(int64_t N, const uint8_t* __restrict__ type, int target,
                                  Layer4<T> dst, T u, T v, T w, T t) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || type[i] != target) return;
    dst.f[0][i] = u;
    dst.f[1][i] = v;
    dst.f[2][i] = w;
    dst.f[3][i] = t;
}

__global__ void zero_f64_kernel(double* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void mask_check_kernel(Geom g, const uint8_t* __restrict__ type,
                                  unsigned long long* __restrict__ out) {
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= g.N) return;
    int k = (int)(id % g.nz);
    int j = (int)((id / g.nz) % g.ny);
    int i = (int)(id / g.sx);
    uint8_t t = type[id];
    bool face = i == 0 || j == 0 || k == 0 || i == g.nx - 1 || j == g.ny - 1 || k == g.nz - 1;
    if (t == NODE_IN && face) atomicAdd(&out[0], 1ull);
    if (t == NODE_OUT) {
        bool adj = (i > 0 && type[id - g.sx] == NODE_IN) || (i < g.nx - 1 && type[id + g.sx] == NODE_IN) ||
                   (j > 0 && type[id - g.sy] == NODE_IN) || (j < g.ny - 1 && type[id + g.sy] == NODE_IN) ||
                   (k > 0 && type[id - 1] == NODE_IN) || (k < g.nz - 1 && type[id + 1] == NODE_IN);
        if (adj) atomicAdd(&out[1], 1ull);
    }
}

}  // namespace

template <class T>
void launch_update_boundaries(const Geom& g, const uint8_t* type, const T* const* gval,
                              Layer4<T> cur, hipStream_t st) {
    hipLaunchKernelGGL((update_boundaries_kernel<T>), dim3(blocks_for(g.N)), dim3(kBS), 0, st, g.N,
                       type, gval[0], gval[1], gval[2], gval[3], cur);
    CMCX_LAUNCH_CHECK();
}

template <class T>
void launch_copy4(Layer4<T> dst, Layer4<T> src, int64_t n, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL((copy4_kernel<T>), dim3(blocks_for(n)), dim3(kBS), 0, st, n, dst, src);
    CMCX_LAUNCH_CHECK();
}

template <class T>
void launch_update_boundaries_list(int64_t n, const uint32_t* ids, const T* const* gval,
                                   Layer4<T> cur, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL((update_boundaries_list_kernel<T>), dim3(blocks_for(n)), dim3(kBS), 0, st, n,
                       ids, gval[0], gval[1], gval[2], gval[3], cur);
    CMCX_LAUNCH_CHECK();
}

template <class T>
void launch_prep_step_list(int64_t n, const uint32_t* ids, const uint8_t* type, Layer4<T> cur,
                           Layer4<T> next, Layer4<T> tA, Layer4<T> tB, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL((prep_step_list_kernel<T>), dim3(blocks_for(n)), dim3(kBS), 0, st, n, ids,
                       type, cur, next, tA, tB);
    CMCX_LAUNCH_CHECK();
}

template <class T>
void launch_prep_step(const Geom& g, const uint8_t* type, Layer4<T> cur, Layer4<T> next,
                      Layer4<T> tA, Layer4<T> tB, hipStream_t st) {
    int full = tB.f[0] == nullptr ? 1 : 0;
    hipLaunchKernelGGL((prep_step_kernel<T>), dim3(blocks_for(g.N)), dim3(kBS), 0, st, g.N, type,
                       cur, next, tA, tB, full);
    CMCX_LAUNCH_CHECK();
}

int div_partials() { return 2 * kDivBlocks; }

template <class T>
void launch_div_error(const Geom& g, const uint8_t* type, Layer4<T> L, T hx, T hy, T hz,
                      double* d_acc, double* d_part, hipStream_t st) {
    const int nb = (int)std::min<int64_t>(kDivBlocks, (g.N + kBS - 1) / kBS);
    hipLaunchKernelGGL((div_error_kernel<T>), dim3(nb), dim3(kBS), 0, st, g, type, L.f[0], L.f[1],
                       L.f[2], hx, hy, hz, d_part);
    CMCX_LAUNCH_CHECK();
    hipLaunchKernelGGL(div_finish_kernel, dim3(1), dim3(64), 0, st, d_part, nb, d_acc);
    CMCX_LAUNCH_CHECK();
}

template <class T>
void launch_scale(const T* src, T* dst, T a, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL((scale_kernel<T>), dim3(blocks_for(n)), dim3(kBS), 0, st, src, dst, a, n);
    CMCX_LAUNCH_CHECK();
}

template <class T>
void launch_convert_from_f64(const double* src, T* dst, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL((convert_kernel<T>), dim3(blocks_for(n)), dim3(kBS), 0, st, src, dst, n);
    CMCX_LAUNCH_CHECK();
}

template <class T>
void launch_copy_typed(const Geom& g, const uint8_t* type, int target, Layer4<T> src,
                       Layer4<T> dst, hipStream_t st) {
    hipLaunchKernelGGL((copy_typed_kernel<T>), dim3(blocks_for(g.N)), dim3(kBS), 0, st, g.N, type,
                       target, src, dst);
    CMCX_LAUNCH_CHECK();
}

template <class T>
void launch_fill_typed(const Geom& g, const uint8_t* type, int target, Layer4<T> dst, T u, T v,
                       T w, T t, hipStream_t st) {
    hipLaunchKernelGGL((fill_typed_kernel<T>), dim3(blocks_for(g.N)), dim3(kBS), 0, st, g.N, type,
                       target, dst, u, v, w, t);
    CMCX_LAUNCH_CHECK();
}

void launch_zero_f64(double* p, int n, hipStream_t st) {
    hipLaunchKernelGGL(zero_f64_kernel, dim3(1), dim3(64), 0, st, p, n);
    CMCX_LAUNCH_CHECK();
}

void launch_mask_check(const Geom& g, const uint8_t* type, unsigned long long* d_out,
                       hipStream_t st) {
    hipLaunchKernelGGL(mask_check_kernel, dim3(blocks_for(g.N)), dim3(kBS), 0, st, g, type, d_out);
    CMCX_LAUNCH_CHECK();
}

#define INST(T)                                                                                   \
    template void launch_update_boundaries<T>(const Geom&, const uint8_t*, const T* const*,        \
                                              Layer4<T>, hipStream_t);                             \
    template void launch_copy4<T>(Layer4<T>, Layer4<T>, int64_t, hipStream_t);                     \
    template void launch_update_boundaries_list<T>(int64_t, const uint32_t*, const T* const*,      \
                                                   Layer4<T>, hipStream_t);                        \
    template void launch_prep_step_list<T>(int64_t, const uint32_t*, const uint8_t*, Layer4<T>,    \
                                           Layer4<T>, Layer4<T>, Layer4<T>, hipStream_t);          \
    template void launch_prep_step<T>(const Geom&, const uint8_t*, Layer4<T>, Layer4<T>,           \
                                      Layer4<T>, Layer4<T>, hipStream_t);                          \
    template void launch_div_error<T>(const Geom&, const uint8_t*, Layer4<T>, T, T, T, double*,    \
                                      double*, hipStream_t);                                       \
    template void launch_scale<T>(const T*, T*, T, int64_t, hipStream_t);                          \
    template void launch_convert_from_f64<T>(const double*, T*, int64_t, hipStream_t);             \
    template void launch_copy_typed<T>(const Geom&, const uint8_t*, int, Layer4<T>, Layer4<T>,     \
                                       hipStream_t);                                               \
    template void launch_fill_typed<T>(const Geom&, const uint8_t*, int, Layer4<T>, T, T, T, T,    \
                                       hipStream_t);
INST(float)
INST(double)
#undef INST

}  // namespace cmcx
