// Internal declarations shared by the HIP translation units of libcmcx.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/cmcx.h"

namespace cmcx {

enum { NODE_IN = 0, NODE_OUT = 1, NODE_BOUND = 2, NODE_VALVE = 3 };
enum { BC_NOSLIP = 0, BC_FREE = 1 };
enum { AX_X = 0, AX_Y = 1, AX_Z = 2 };

void set_error(const std::string& msg);

struct HipError {
    hipError_t code;
    std::string where;
};

#define CMCX_HIP(expr)                                                            \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) throw ::cmcx::HipError{_e, std::string(#expr)};     \
    } while (0)

#define CMCX_LAUNCH_CHECK()                                                       \
    do {                                                                          \
        hipError_t _e = hipGetLastError();                                        \
        if (_e != hipSuccess) throw ::cmcx::HipError{_e, std::string(__func__)};  \
    } while (0)

// Geometry of one grid (dimensions and strides of the x-major, z-fastest layout).
struct Geom {
    int nx, ny, nz;
    int64_t N;
    int64_t sx, sy;  // strides of i and j (sz = 1)
};

// Segment lists of one axis as structure-of-arrays on the device.
struct SegList {
    int64_t n = 0;
    int max_size = 0;
    int64_t sum_size = 0;      // S_d: sum of segment sizes, endpoints included
    int32_t* d = nullptr;      // 7 arrays of n int32: px, py, pz, ex, ey, ez, size
};

// Field-type dependent parameters of one sweep, precomputed on the host in the
// field type T exactly as the reference computes them (AdiSolver3D.cpp:739-753).
template <class T>
struct SweepParams {
    T dt;
    T h;          // spacing along the sweep axis
    T hx, hy, hz; // spacings (stencils of the dissipation terms)
    T vis_v;      // v_vis / h^2
    T vis_t;      // t_vis / h^2
    T v_T, t_phi;
    // reciprocals for the fused kernels (multiplications instead of divisions)
    T three_dt;   // 3 / dt
    T inv_2h;     // 1 / (2 h)
    T inv_2hx, inv_2hy, inv_2hz;
};

template <class T>
struct Layer4 {
    T* f[4];
};

// Segment builder (kernels_segments.hip)
void build_segment_list(const uint8_t* d_type, const Geom& g, int axis, int nblockz,
                        SegList& out, hipStream_t st);

// Plain sweep (kernels_plain.hip)
template <class T>
void launch_plain_sweep(const Geom& g, int axis, const SegList& segs, const SweepParams<T>& p,
                        const uint8_t* type, const uint8_t* bcv, const uint8_t* bct, const T* const* gval,
                        Layer4<T> in, Layer4<T> tin, Layer4<T> out, Layer4<T> tout, int merges,
                        int var0, int var1, T* scratch, hipStream_t st);

// Fused sweep (kernels_fused.hip).  A launch covers a batch of independent
// grids stored back to back in the flat arrays; `variant` selects, for the
// strided (X, Y) sweeps, the tile width (0 = 128-byte rows, 1 = 64-byte rows,
// 2 = 32-byte rows, each with chunks of 8 rows per lane; 3-5 the same with
// chunks of 4 rows, for lines of at most 128 nodes; 6-9 = variants 1, 2, 4, 5
// with the Picard layer kept in LDS for the merge; strided_lpg gives the z
// lines of a tile) and, for the Z sweep, bit 0 = registers capped for more
// waves per SIMD (else the compiler's choice), bit 1 = the Picard layer parked
// in LDS until the merge, bit 2 = a line spread over 64 lanes (else 32).
constexpr int kFusedVariants = 10;  // strided sweeps
constexpr int kZVariants = 8;
constexpr int kZWaves = 2;  // wavefronts per Z-sweep work-group (two lines each)

// Descriptor of one grid of a batch.  three_dt is passed per launch.
template <class T>
struct InstDesc {
    Geom g;
    int64_t off;            // flat index of node (0, 0, 0)
    SweepParams<T> p[3];    // per sweep axis
};

// Work-group list of one fused sweep over all grids of a batch.
struct FusedWork {
    int64_t nblocks = 0;
    int32_t* d_items = nullptr;  // X/Y: {grid, kmax tiles} per work-group; Z: 2*kZWaves
                                 // slots (first line x*ny + y, line count) per work-group
    int32_t* d_inst = nullptr;   // Z: grid of each work-group
    int maxn = 0;                // longest line along the axis over the grids
    int kmax = 1;                // X/Y: tiles per work item
    int vec_ok = 1;              // Z: every line starts on a 16-byte boundary
    int nz_gcd = 0;              // Z: greatest common divisor of the grids' nz
    int stream = 0;              // sweeps store past the caches (kStoreStream)
    void release();
};

// Stores bit of the fused sweeps: non-temporal stores of the output and the
// merged Picard layer.  Set for launches whose 16 fields moved per sweep
// exceed kStreamStoreBytes, the MI355X memory-side cache (256 MiB).  Measured
// on the ellipsoids workload, FP32, step time with streamed against plain
// stores: 128^3 (134 MB per sweep, which the cache holds) +14 %, 160^3
// (262 MB) -0.8 %, 192^3 (453 MB) -6.2 %, 256^3 (1.07 GB) -8 to -10 %.
constexpr int kStoreStream = 4;
constexpr int64_t kStreamStoreBytes = 256ll << 20;

template <class T>
bool fused_supported(const Geom& g);
template <class T>
int strided_lpg(int variant);
template <class T>
int strided_chunk_rows(int variant);
// Strided variant v usable for lines of n nodes.
template <class T>
bool fused_variant_ok(int n, int v);
// Node flags of the fused kernels: type | kFlagFreeVel | kFlagFreeT (0 on NODE_IN).
constexpr uint8_t kFlagFreeVel = 4, kFlagFreeT = 8;
void build_node_flags(const uint8_t* t, const uint8_t* bv, const uint8_t* bt, int64_t n,
                      uint8_t* out);
// bsync: the state holds the node values on BOUND/VALVE nodes with no-slip BCs
// (the Picard layer then supplies the Dirichlet values).  stores: bit 0 writes
// the sweep output, bit 1 the merged Picard layer (3 = both).
template <class T>
void launch_fused_sweep(const InstDesc<T>* insts, T three_dt, int axis, const FusedWork& w,
                        const uint8_t* tflags, const T* const* gval, Layer4<T> in, Layer4<T> tin,
                        Layer4<T> out, Layer4<T> tout, int merges, int bsync, int stores,
                        int variant, hipStream_t st);
// Builds the work of `axis` from the host copy of the flat node types; lpg and
// m: z lines per tile and rows per chunk of the strided (X, Y) sweeps; elem:
// sizeof the field type.  Tiles of grids with lines shorter than the longest
// are packed several to a work-group (length binning inside one launch).
void build_fused_work(const uint8_t* h_type, const std::vector<Geom>& geoms,
                      const std::vector<int64_t>& offs, int axis, int lpg, int m, size_t elem,
                      FusedWork& w);
// Axes whose lines hold a BC_FREE node with NODE_IN on both sides (a free node
// shared by two segments): bit a set for axis a.  Host arrays of one grid.
int shared_free_axes(const uint8_t* type, const uint8_t* bcv, const uint8_t* bct, const Geom& g);

// Adjoint sweeps (kernels_adjoint.hip).  Lines per tile of the transposed line
// solve (adj_lpg: 128 or 64 bytes of lines); every axis at most 256 nodes long.
template <class T>
int adj_lpg(const Geom& g);
template <class T>
bool adjoint_supported(const Geom& g);
// Tiles (outer * ltiles + line tile) of `axis` holding a NODE_IN node, from host types.
void build_adj_tiles(const uint8_t* h_type, const Geom& g, int axis, int lpg, int32_t** d_tiles,
                     int64_t* ntiles, int* ltiles);
// Number of per-block parameter partial slots (4 doubles each) of launch_adj_sweep.
int adj_point_blocks();
// Reverses one sweep: xscale * xs (adjoint of the sweep output; xs.f[0] null
// for zero; xs may be `lam`, read before it is overwritten) and tb (adjoint of
// the Picard layer after the sweep) -> lam, tb (before it); the input adjoint
// is (3/dt) lam on NODE_IN nodes, left in lam for the next reversed sweep, or
// with zmode (the input is the state) added to cacc (first_z: cacc set);
// gb += lam on the node-value rows; pp[4*b..] += parameter partials (v_T,
// v_vis, t_vis, t_phi).  `last` (the Z sweep of the first Picard iteration)
// also ends the reversed step: xb = state adjoint, gb += tbar on BOUND/VALVE
// nodes, tb cleared.
template <class T>
void launch_adj_sweep(const Geom& g, int axis, const SweepParams<T>& p, const uint8_t* type,
                      const uint8_t* bcv, const uint8_t* bct, Layer4<T> tin, Layer4<T> out,
                      Layer4<T> xs, T xscale, Layer4<T> xb, Layer4<T> tb, Layer4<T> cacc,
                      Layer4<T> lam, Layer4<T> gb, T alpha, T beta, int zmode, int first_z, int last,
                      const int32_t* tiles, int64_t ntiles, int ltiles, double* pp,
                      const uint16_t* nmask, hipStream_t st);
// Node words of the pointwise adjoint kernel (type, free BC kinds, NODE_IN
// neighbours), built once per backward pass: N 16-bit words.
void launch_adj_mask(const Geom& g, const uint8_t* type, const uint8_t* bcv, const uint8_t* bct,
                     uint16_t* mask, hipStream_t st);
// dst <- src, four fields of n elements, one launch.
template <class T>
void launch_copy4(Layer4<T> dst, Layer4<T> src, int64_t n, hipStream_t st);

// Field kernels (kernels_fields.hip)
template <class T>
void launch_update_boundaries(const Geom& g, const uint8_t* type, const T* const* gval,
                              Layer4<T> cur, hipStream_t st);
template <class T>
void launch_update_boundaries_list(int64_t n, const uint32_t* ids, const T* const* gval,
                                   Layer4<T> cur, hipStream_t st);
template <class T>
void launch_prep_step_list(int64_t n, const uint32_t* ids, const uint8_t* type, Layer4<T> cur,
                           Layer4<T> next, Layer4<T> tA, Layer4<T> tB, hipStream_t st);
template <class T>
void launch_prep_step(const Geom& g, const uint8_t* type, Layer4<T> cur, Layer4<T> next,
                      Layer4<T> tA, Layer4<T> tB, hipStream_t st);
// Number of doubles of the per-block partials buffer of launch_div_error.
int div_partials();
template <class T>
void launch_div_error(const Geom& g, const uint8_t* type, Layer4<T> L, T hx, T hy, T hz,
                      double* d_acc, double* d_part, hipStream_t st);
template <class T>
void launch_scale(const T* src, T* dst, T a, int64_t n, hipStream_t st);
template <class T>
void launch_convert_from_f64(const double* src, T* dst, int64_t n, hipStream_t st);
template <class T>
void launch_copy_typed(const Geom& g, const uint8_t* type, int target, Layer4<T> src,
                       Layer4<T> dst, hipStream_t st);
template <class T>
void launch_fill_typed(const Geom& g, const uint8_t* type, int target, Layer4<T> dst, T u, T v,
                       T w, T t, hipStream_t st);
void launch_zero_f64(double* p, int n, hipStream_t st);
// mask checks: out[0] = #IN nodes on box faces, out[1] = #OUT nodes 6-adjacent to IN
void launch_mask_check(const Geom& g, const uint8_t* type, unsigned long long* d_out,
                       hipStream_t st);

}  // namespace cmcx
