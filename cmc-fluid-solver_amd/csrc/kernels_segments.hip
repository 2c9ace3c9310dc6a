// Segment builder: cuts every X, Y and Z grid line into tridiagonal systems.
//
// Semantics of Grid3D::GenerateListSegments (src/FluidSolver3D/Grid3D.cpp:47-127),
// rebuilt as three device passes (count per line, exclusive scan, emit) so the
// list comes out in the reference's order without a host scan of the mask:
//   * a segment starts at the non-IN node before a run of NODE_IN nodes and ends
//     at the first non-IN node after it (size = run + 2);
//   * node 0 of a line is never tested, a run reaching the last node is dropped;
//   * line order X: (y outer, z inner), Y: (x, z), Z: (x, y) -- AdiSolver3D.cpp:557-559;
//   * nblockz > 1 regroups the X and Y lines by z blocks (Grid3D.cpp:59-62, 124-125).
#include "cmcx_internal.h"

#include <algorithm>

namespace cmcx {

namespace {

constexpr int kMaxBlocksZ = 64;

struct LineMap {
    int axis;
    int dim2, dim3;          // outer and inner line coordinates
    int nb;                  // number of z blocks
    int bs[kMaxBlocksZ + 1]; // block boundaries along dim3
    int64_t cum[kMaxBlocksZ + 1];  // lines before block b
};

__device__ inline void line_start(const LineMap& m, const Geom& g, int64_t L, int& px, int& py,
                                  int& pz) {
    int b = 0;
    while (b + 1 < m.nb && L >= m.cum[b + 1]) b++;
    int64_t loc = L - m.cum[b];
    int w = m.bs[b + 1] - m.bs[b];
    int i = (int)(loc / w);
    int j = m.bs[b] + (int)(loc % w);
    if (m.axis == AX_X) { px = 0; py = i; pz = j; }
    else if (m.axis == AX_Y) { px = i; py = 0; pz = j; }
    else { px = i; py = j; pz = 0; }
}

__global__ void seg_count_kernel(const uint8_t* __restrict__ type, Geom g, LineMap m, int64_t nlines,
                                 int32_t* __restrict__ counts,
                                 unsigned long long* __restrict__ stats) {
    int64_t L = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= nlines) return;
    int px, py, pz;
    line_start(m, g, L, px, py, pz);
    int dim1 = m.axis == AX_X ? g.nx : m.axis == AX_Y ? g.ny : g.nz;
    int64_t stride = m.axis == AX_X ? g.sx : m.axis == AX_Y ? g.sy : 1;
    int64_t base = px * g.sx + py * g.sy + pz;
    int cnt = 0, state = 0, start = 0;
    unsigned long long ssum = 0, smax = 0;
    for (int t = 0; t + 1 < dim1; t++) {
        bool in = type[base + (int64_t)(t + 1) * stride] == NODE_IN;
        if (in) {
            if (state == 0) start = t;
            state = 1;
        } else if (state == 1) {
            int size = (t + 1) - start + 1;
            cnt++;
            ssum += size;
            smax = size > (int)smax ? size : smax;
            state = 0;
        }
    }
    counts[L] = cnt;
    if (ssum) {
        atomicAdd(&stats[0], ssum);
        atomicMax(&stats[1], smax);
    }
}

// Exclusive scan of counts (one block): off[0..n], off[n] = total.
__global__ void scan_kernel(const int32_t* __restrict__ counts, int64_t n, int64_t* __restrict__ off) {
    __shared__ int64_t part[1024];
    int tid = threadIdx.x;
    int64_t chunk = (n + blockDim.x - 1) / blockDim.x;
    int64_t lo = tid * chunk, hi = lo + chunk < n ? lo + chunk : n;
    int64_t s = 0;
    for (int64_t i = lo; i < hi; i++) s += counts[i];
    part[tid] = s;
    __syncthreads();
    for (int d = 1; d < (int)blockDim.x; d <<= 1) {
        int64_t v = tid >= d ? part[tid - d] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int64_t run = tid ? part[tid - 1] : 0;
    for (int64_t i = lo; i < hi; i++) {
        off[i] = run;
        run += counts[i];
    }
    if (tid == (int)blockDim.x - 1) off[n] = part[tid];
}

__global__ void seg_emit_kernel(const uint8_t* __restrict__ type, Geom g, LineMap m, int64_t nlines,
                                const int64_t* __restrict__ off, int32_t* __restrict__ segs,
                                int64_t nseg) {
    int64_t L = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= nlines) return;
    int px, py, pz;
    line_start(m, g, L, px, py, pz);
    int dim1 = m.axis == AX_X ? g.nx : m.axis == AX_Y ? g.ny : g.nz;
    int64_t stride = m.axis == AX_X ? g.sx : m.axis == AX_Y ? g.sy : 1;
    int64_t base = px * g.sx + py * g.sy + pz;
    int64_t o = off[L];
    int state = 0, start = 0;
    for (int t = 0; t + 1 < dim1; t++) {
        bool in = type[base + (int64_t)(t + 1) * stride] == NODE_IN;
        if (in) {
            if (state == 0) start = t;
            state = 1;
        } else if (state == 1) {
            int e = t + 1;
            int32_t sx0 = px, sy0 = py, sz0 = pz, ex = px, ey = py, ez = pz;
            if (m.axis == AX_X) { sx0 = start; ex = e; }
            else if (m.axis == AX_Y) { sy0 = start; ey = e; }
            else { sz0 = start; ez = e; }
            segs[0 * nseg + o] = sx0;
            segs[1 * nseg + o] = sy0;
            segs[2 * nseg + o] = sz0;
            segs[3 * nseg + o] = ex;
            segs[4 * nseg + o] = ey;
            segs[5 * nseg + o] = ez;
            segs[6 * nseg + o] = e - start + 1;
            o++;
            state = 0;
        }
    }
}

LineMap make_map(const Geom& g, int axis, int nblockz) {
    LineMap m{};
    m.axis = axis;
    if (axis == AX_X) { m.dim2 = g.ny; m.dim3 = g.nz; }
    else if (axis == AX_Y) { m.dim2 = g.nx; m.dim3 = g.nz; }
    else { m.dim2 = g.nx; m.dim3 = g.ny; }
    int nb = axis == AX_Z ? 1 : std::max(1, nblockz);
    // block boundaries exactly as the reference advances block_end (Grid3D.cpp:60-62,124-125)
    int d3 = m.dim3, q = d3 / nb;
    int bstart = 0, bend = q, k = 0;
    m.bs[0] = 0;
    while (bend <= d3 && k < kMaxBlocksZ) {
        m.bs[k + 1] = bend;
        k++;
        bstart = bend;
        bend += (bend + q == q * nb) ? q + d3 % nb : q;
    }
    (void)bstart;
    m.nb = k;
    m.cum[0] = 0;
    for (int b = 0; b < k; b++) m.cum[b + 1] = m.cum[b] + (int64_t)m.dim2 * (m.bs[b + 1] - m.bs[b]);
    return m;
}

}  // namespace

void build_segment_list(const uint8_t* d_type, const Geom& g, int axis, int nblockz, SegList& out,
                        hipStream_t st) {
    if (nblockz > 1 && axis != AX_Z) {
        int d3 = g.nz;
        if (nblockz > d3 || nblockz > kMaxBlocksZ) throw std::string("nblockz out of range");
    }
    LineMap m = make_map(g, axis, nblockz);
    int64_t nlines = m.cum[m.nb];
    int32_t* d_counts = nullptr;
    int64_t* d_off = nullptr;
    unsigned long long* d_stats = nullptr;
    CMCX_HIP(hipMallocAsync((void**)&d_counts, sizeof(int32_t) * (nlines + 1), st));
    CMCX_HIP(hipMallocAsync((void**)&d_off, sizeof(int64_t) * (nlines + 1), st));
    CMCX_HIP(hipMallocAsync((void**)&d_stats, sizeof(unsigned long long) * 2, st));
    CMCX_HIP(hipMemsetAsync(d_stats, 0, sizeof(unsigned long long) * 2, st));
    int bs = 256;
    int64_t nb = (nlines + bs - 1) / bs;
    hipLaunchKernelGGL(seg_count_kernel, dim3((unsigned)nb), dim3(bs), 0, st, d_type, g, m, nlines,
                       d_counts, d_stats);
    CMCX_LAUNCH_CHECK();
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, st, d_counts, nlines, d_off);
    CMCX_LAUNCH_CHECK();
    int64_t total = 0;
    unsigned long long stats[2] = {0, 0};
    CMCX_HIP(hipMemcpyAsync(&total, d_off + nlines, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    CMCX_HIP(hipMemcpyAsync(stats, d_stats, sizeof(stats), hipMemcpyDeviceToHost, st));
    CMCX_HIP(hipStreamSynchronize(st));
    if (out.d) CMCX_HIP(hipFree(out.d));
    out.d = nullptr;
    out.n = total;
    out.sum_size = (int64_t)stats[0];
    out.max_size = (int)stats[1];
    if (total > 0) {
        CMCX_HIP(hipMalloc((void**)&out.d, sizeof(int32_t) * 7 * total));
        hipLaunchKernelGGL(seg_emit_kernel, dim3((unsigned)nb), dim3(bs), 0, st, d_type, g, m,
                           nlines, d_off, out.d, total);
        CMCX_LAUNCH_CHECK();
    }
    CMCX_HIP(hipFreeAsync(d_counts, st));
    CMCX_HIP(hipFreeAsync(d_off, st));
    CMCX_HIP(hipFreeAsync(d_stats, st));
    CMCX_HIP(hipStreamSynchronize(st));
}

}  // namespace cmcx
