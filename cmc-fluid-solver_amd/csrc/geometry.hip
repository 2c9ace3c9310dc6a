// Geometry ingestion: the reference's Shape3D triangle meshes and Shape2D
// polylines extruded in depth (src/FluidSolver3D/Grid3D.cpp:373-436, 488-514,
// 608-666, 690-912; src/FluidSolver2D/Grid2D.cpp:105-356) turned into the node
// arrays of a 3D grid (type, BC kinds, node velocity and temperature) on the
// GPU, for cmcx_create.
//
// Files are parsed on the host; a build blends the frames at the requested
// time (as the reference's Prepare) and then, on the device:
//   triangle mesh: every node NODE_IN; one thread per triangle rasterises it
//     (scanline over the triangle projected on the plane of its largest normal
//     component, each point projected back onto the triangle's plane) and its
//     three edges as NODE_BOUND; an iterated flood fill from node (0, 0, 0)
//     through NODE_IN nodes marks the outside NODE_OUT; NODE_IN and NODE_OUT
//     nodes get no-slip BCs, v = 0, T = baseT (walls keep what they hold: 0).
//   extruded polylines: a 2D grid of cells, one thread per polyline segment
//     (valve segments of moving shapes first, then walls); a cell several
//     segments or steps of one segment reach takes the last of them in the
//     reference's sequential order, found with a 64-bit (order, step) key per
//     cell (atomic maximum, then the owner writes); a flood fill from cell
//     (0, 0); then one thread per column lays the column out in depth (outside,
//     bottom wall whose height follows the depth variation, fluid or the 2D
//     cell's wall or valve, top wall, outside).
// Every rasterised point, order of arithmetic and float rounding follows the
// reference's single-precision code (FP contraction off), so the arrays are
// bit-identical to a sequential build.  The reference's scanline loops have no
// bounds; here every loop is capped by the grid's extent and out-of-grid
// points are skipped (a well-formed input never reaches either).
//
// Input text is read as the reference reads it on its own platform: Windows
// text mode, where a CRLF line end arrives as LF.
#include "cmcx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace cmcx {

namespace {

constexpr float kScale = 0.001f;  // GRID_SCALE_FACTOR: file units (mm) to metres
constexpr double kPad = 0.02;     // BBOX_PADDING
constexpr double kEps = 1e-8;     // COMP_EPS

struct F2 {
    float x, y;
};
struct F3 {
    float x, y, z;
};

int align32(int n) { return (n & 31) == 0 ? n : ((n >> 5) + 1) << 5; }

// ---- text input -------------------------------------------------------------
struct Text {
    std::string s;
    size_t i = 0;
    bool load(const char* path) {
        std::ifstream f(path, std::ios::binary);
        if (!f) return false;
        std::stringstream ss;
        ss << f.rdbuf();
        for (char c : ss.str())
            if (c != '\r') s.push_back(c);  // text mode: CRLF arrives as LF
        return true;
    }
    void skip_ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\t')) i++;
    }
    bool integer(int& v) {
        skip_ws();
        size_t j = i;
        if (j < s.size() && (s[j] == '-' || s[j] == '+')) j++;
        while (j < s.size() && isdigit((unsigned char)s[j])) j++;
        if (j == i) return false;
        v = atoi(s.substr(i, j - i).c_str());
        i = j;
        return true;
    }
    std::string word() {
        skip_ws();
        size_t j = i;
        while (j < s.size() && s[j] != ' ' && s[j] != '\n' && s[j] != '\t') j++;
        std::string w = s.substr(i, j - i);
        i = j;
        return w;
    }
    // the rest of the current line after leading blanks (ReadPoint2D)
    std::string line() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\n')) i++;
        size_t j = s.find('\n', i);
        if (j == std::string::npos) j = s.size();
        std::string l = s.substr(i, j - i);
        i = j;
        return l;
    }
};

float to_float(std::string t) {
    for (char& c : t)
        if (c == ',') c = '.';
    return (float)atof(t.c_str());
}

// ReadPoint2D: a line of two values separated by a space (',' decimal marks
// allowed)
F2 point2(Text& tx) {
    std::string l = tx.line();
    const size_t k = l.find(' ');
    return F2{to_float(l.substr(0, k)), to_float(k == std::string::npos ? "" : l.substr(k + 1))};
}

// ---- device helpers (single precision in the reference's operation order) ----
#pragma clang fp contract(off)

__device__ __forceinline__ F2 horizon(F2 p1, F2 p2, F2 p) {
    F2 r;
    r.y = p.y;
    if (fabs((double)(p1.y - p2.y)) < kEps) r.x = p.x;
    else r.x = p1.x + (p2.x - p1.x) * (r.y - p1.y) / (p2.y - p1.y);
    return r;
}

struct Box {
    int nx, ny, nz;
    __device__ bool in(int i, int j, int k) const {
        return i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz;
    }
    __device__ int64_t id(int i, int j, int k) const { return ((int64_t)i * ny + j) * nz + k; }
};

__device__ __forceinline__ void mark(uint8_t* type, const Box& b, int i, int j, int k) {
    if (b.in(i, j, k)) type[b.id(i, j, k)] = NODE_BOUND;
}

// ProjectPointOnPolygon: the node of the plane point above (i, j) on the axis
// `dir` of the largest normal component
__device__ __forceinline__ void project(uint8_t* type, const Box& b, int dir, int i, int j, float py,
                                        F3 n, float d) {
    const float ti = (float)i;
    if (dir == 0) {
        const int k = (int)((-d - (ti * n.y + py * n.z)) / n.x);
        if (k >= 0 && k < b.nx) mark(type, b, k, i, j);
    } else if (dir == 1) {
        const int k = (int)((-d - (ti * n.x + py * n.z)) / n.y);
        if (k >= 0 && k < b.ny) mark(type, b, i, k, j);
    } else {
        const int k = (int)((-d - (ti * n.x + py * n.y)) / n.z);
        if (k >= 0 && k < b.nz) mark(type, b, i, j, k);
    }
}

// One half of the triangle's scanline: rows from p (advancing by dp) while
// p.y < yend, each row from p.x to the edge (e1, e2).
__device__ __forceinline__ void scan_half(uint8_t* type, const Box& b, int dir, F3 n, float d, F2& p,
                                          F2 dp, float yend, F2 e1, F2 e2, int di, int cap) {
    for (int guard = 0; p.y < yend && guard < cap; guard++) {
        const int j = (int)p.y;
        const int last = (int)horizon(e1, e2, p).x;
        int i = (int)p.x;
        for (int g2 = 0; i != last + di && g2 < cap; g2++, i += di) project(type, b, dir, i, j, p.y, n, d);
        p.x += dp.x;
        p.y += dp.y;
    }
}

// RasterPolygon and the three RasterLine calls of one triangle (Grid3D.cpp:874-886)
__global__ void raster_tri_kernel(const F3* __restrict__ v, const int* __restrict__ idx, int ntri,
                                  Box b, uint8_t* __restrict__ type) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    const F3 p1 = v[idx[3 * t]], p2 = v[idx[3 * t + 1]], p3 = v[idx[3 * t + 2]];
    const int cap = 4 * (b.nx + b.ny + b.nz) + 16;  // longer than any in-grid scan
    const bool degenerate =
        fabs((double)(p1.x - p2.x)) < kEps && fabs((double)(p1.y - p2.y)) < kEps &&
        fabs((double)(p1.z - p2.z)) < kEps && fabs((double)(p1.x - p3.x)) < kEps &&
        fabs((double)(p1.y - p3.y)) < kEps && fabs((double)(p1.z - p3.z)) < kEps;
    if (!degenerate) {
        const F3 e1{p2.x - p1.x, p2.y - p1.y, p2.z - p1.z}, e2{p3.x - p1.x, p3.y - p1.y, p3.z - p1.z};
        F3 n{e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
        const float inv = 1 / sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
        n.x *= inv;
        n.y *= inv;
        n.z *= inv;
        const float d = -(p1.x * n.x + p1.y * n.y + p1.z * n.z);
        const float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
        const float mx = fmaxf(ax, fmaxf(ay, az));
        int dir = 0;
        if (fabs((double)(mx - ax)) < kEps) dir = 0;
        if (fabs((double)(mx - ay)) < kEps) dir = 1;
        if (fabs((double)(mx - az)) < kEps) dir = 2;
        F2 a, c, e;
        if (dir == 0) { a = {p1.y, p1.z}; c = {p2.y, p2.z}; e = {p3.y, p3.z}; }
        else if (dir == 1) { a = {p1.x, p1.z}; c = {p2.x, p2.z}; e = {p3.x, p3.z}; }
        else { a = {p1.x, p1.y}; c = {p2.x, p2.y}; e = {p3.x, p3.y}; }
        // sort by y: a <= c <= e
        F2 m;
        if (e.y < c.y) { m = e; e = c; c = m; }
        if (a.y > c.y) { m = a; a = c; c = m; }
        if (e.y < c.y) { m = e; e = c; c = m; }
        m = horizon(a, e, c);
        const F2 d1{m.x - a.x, m.y - a.y}, d2{e.x - m.x, e.y - m.y};
        const int s1 = (int)fmaxf(fabsf(d1.x), fabsf(d1.y)) + 1;
        const int s2 = (int)fmaxf(fabsf(d2.x), fabsf(d2.y)) + 1;
        const F2 dp1{d1.x / s1, d1.y / s1}, dp2{d2.x / s2, d2.y / s2};
        F2 p = a;
        const int di = m.x < c.x ? 1 : -1;
        scan_half(type, b, dir, n, d, p, dp1, m.y, a, c, di, cap);
        scan_half(type, b, dir, n, d, p, dp2, e.y, c, e, di, cap);
    }
    // edges (RasterLine, Grid3D.cpp:803-821): p1-p2, p1-p3, p3-p2
    const F3 ends[3][2] = {{p1, p2}, {p1, p3}, {p3, p2}};
    for (int q = 0; q < 3; q++) {
        const F3 s0 = ends[q][0], s1e = ends[q][1];
        const F3 dir{s1e.x - s0.x, s1e.y - s0.y, s1e.z - s0.z};
        const int steps = (int)fmaxf(fabsf(dir.x), fmaxf(fabsf(dir.y), fabsf(dir.z))) + 1;
        const float s = 1 / (float)steps;
        const F3 dp{dir.x * s, dir.y * s, dir.z * s};
        F3 p = s0;
        for (int i = 0; i <= steps && i <= cap; i++) {
            mark(type, b, (int)p.x, (int)p.y, (int)p.z);
            p.x += dp.x;
            p.y += dp.y;
            p.z += dp.z;
        }
    }
}

#pragma clang fp contract(on)

__global__ void fill_u8_kernel(uint8_t* __restrict__ a, int64_t n, uint8_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// One flood-fill round: NODE_IN nodes with a NODE_OUT face neighbour become
// NODE_OUT (in place, so a round can carry the front further); *changed is set
// when a node changed.  nz = 1 for a 2D grid (4 neighbours).
__global__ void flood_kernel(uint8_t* __restrict__ type, Box b, int* __restrict__ changed) {
    const int64_t n = (int64_t)b.nx * b.ny * b.nz;
    int any = 0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        if (type[q] != NODE_IN) continue;
        const int k = (int)(q % b.nz);
        const int64_t l = q / b.nz;
        const int j = (int)(l % b.ny), i = (int)(l / b.ny);
        const int64_t sx = (int64_t)b.ny * b.nz, sy = b.nz;
        bool out = (i > 0 && type[q - sx] == NODE_OUT) || (i + 1 < b.nx && type[q + sx] == NODE_OUT) ||
                   (j > 0 && type[q - sy] == NODE_OUT) || (j + 1 < b.ny && type[q + sy] == NODE_OUT) ||
                   (k > 0 && type[q - 1] == NODE_OUT) || (k + 1 < b.nz && type[q + 1] == NODE_OUT);
        if (out) {
            type[q] = NODE_OUT;
            any = 1;
        }
    }
    if (any) *changed = 1;
}

// Build's final pass (Grid3D.cpp:898-911): fluid and outside nodes get no-slip
// BCs, v = 0 and T = baseT; walls keep their data.
__global__ void mesh_data_kernel(const uint8_t* __restrict__ type, int64_t n, uint8_t* __restrict__ bcv,
                                 uint8_t* __restrict__ bct, float* __restrict__ v0,
                                 float* __restrict__ v1, float* __restrict__ v2,
                                 float* __restrict__ T, float base_t) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const uint8_t t = type[q];
    if (t == NODE_IN || t == NODE_OUT) {
        bcv[q] = BC_NOSLIP;
        bct[q] = BC_NOSLIP;
        v0[q] = v1[q] = v2[q] = 0.0f;
        T[q] = base_t;
    }
}

// ---- 2D polylines -------------------------------------------------------------
struct Seg2 {
    F2 p1, p2, v1, v2;
    int color;
};

#pragma clang fp contract(off)

// RasterLine (Grid2D.cpp:105-141): the cells of segment s, step by step;
// mode 0 claims each cell with the key (s, step), mode 1 writes the cells the
// segment still owns.  Cells take (no-slip, color, v, startT).
__global__ void raster_seg2_kernel(const Seg2* __restrict__ segs, int nseg, int nx, int ny, int mode,
                                   unsigned long long* __restrict__ key, uint8_t* __restrict__ cell,
                                   float* __restrict__ vx, float* __restrict__ vy,
                                   float* __restrict__ T, float start_t) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    const Seg2 g = segs[s];
    const F2 o{g.p2.x - g.p1.x, g.p2.y - g.p1.y};
    const int steps = (int)fmaxf(fabsf(o.x), fabsf(o.y)) + 1;
    const F2 dp{o.x / steps, o.y / steps};
    const F2 dv{(g.v2.x - g.v1.x) / steps, (g.v2.y - g.v1.y) / steps};
    F2 p = g.p1, v = g.v1;
    const int cap = 4 * (nx + ny) + 16;
    for (int i = 0; i <= steps && i <= cap; i++) {
        const int x = (int)p.x, y = (int)p.y;
        if (x >= 0 && x < nx && y >= 0 && y < ny) {
            const int64_t c = (int64_t)x * ny + y;
            const unsigned long long k = ((unsigned long long)s << 32) | (unsigned)i;
            if (mode == 0) {
                atomicMax(&key[c], k);
            } else if (key[c] == k) {
                cell[c] = (uint8_t)g.color;
                vx[c] = v.x;
                vy[c] = v.y;
                T[c] = start_t;
            }
        }
        p.x += dp.x;
        p.y += dp.y;
        v.x += dv.x;
        v.y += dv.y;
    }
}

#pragma clang fp contract(on)

// Grid2D::Build's final pass: fluid and outside cells get v = 0, T = startT
// (the BC kind of every cell is no-slip)
__global__ void cell_data_kernel(const uint8_t* __restrict__ cell, int64_t n, float* __restrict__ vx,
                                 float* __restrict__ vy, float* __restrict__ T, float start_t) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    if (cell[c] == NODE_IN || cell[c] == NODE_OUT) {
        vx[c] = vy[c] = 0.0f;
        T[c] = start_t;
    }
}

// Prepare2D (Grid3D.cpp:608-666): one thread per (i, j) column.
__global__ void extrude_kernel(const uint8_t* __restrict__ cell, const float* __restrict__ cvx,
                               const float* __restrict__ cvy, const float* __restrict__ cT, Box b,
                               int active_nz, double depth_var, float base_t,
                               uint8_t* __restrict__ type, uint8_t* __restrict__ bcv,
                               uint8_t* __restrict__ bct, float* __restrict__ v0,
                               float* __restrict__ v1, float* __restrict__ v2,
                               float* __restrict__ T) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= b.nx * b.ny) return;
    const int i = col / b.ny, j = col % b.ny;
    const int64_t base = (int64_t)col * b.nz;
    auto set = [&](int k, uint8_t t, uint8_t bv, uint8_t bt, float ux, float uy, float tt) {
        const int64_t q = base + k;
        type[q] = t;
        bcv[q] = bv;
        bct[q] = bt;
        v0[q] = ux;
        v1[q] = uy;
        v2[q] = 0.0f;
        T[q] = tt;
    };
    // the column starts zeroed (NODE_IN, no-slip, v = 0, T = 0)
    for (int k = 0; k < b.nz; k++) set(k, NODE_IN, BC_NOSLIP, BC_NOSLIP, 0.0f, 0.0f, 0.0f);
    const uint8_t c = cell[col];
    if (c == NODE_OUT) {
        for (int k = 0; k < b.nz; k++) type[base + k] = NODE_OUT;
        return;
    }
    for (int k = active_nz - 1; k < b.nz; k++) type[base + k] = NODE_OUT;
    if (active_nz - 2 >= 0) set(active_nz - 2, NODE_BOUND, BC_NOSLIP, BC_FREE, 0.0f, 0.0f, base_t);
    const int height = max(active_nz - 2 - 2, 0);
    const double x = -1 + 2 * (double)i / b.nx;
    const double y = -1 + 2 * (double)j / b.ny;
    const double z = 1.0 - (x * x + y * y) * 0.5;
    const int bottom = 1 + (int)(depth_var * z * height);
    type[base] = NODE_OUT;
    for (int k = 1; k <= bottom && k < b.nz; k++) set(k, NODE_BOUND, BC_NOSLIP, BC_FREE, 0.0f, 0.0f, base_t);
    const float ux = cvx[col], uy = cvy[col], ct = cT[col];
    for (int k = bottom + 1; k < active_nz - 2; k++) {
        if (c == NODE_BOUND) {
            set(k, NODE_BOUND, BC_NOSLIP, BC_FREE, ux, uy, ct);
        } else if (c == NODE_VALVE) {
            if (ux == 0 && uy == 0) set(k, NODE_VALVE, BC_FREE, BC_FREE, ux, uy, ct);
            else set(k, NODE_VALVE, BC_NOSLIP, BC_NOSLIP, ux, uy, ct);
        } else {
            type[base + k] = NODE_IN;
            T[base + k] = base_t;
        }
    }
}

// ---- host objects -------------------------------------------------------------
struct Frame3 {
    std::vector<F3> v;
    std::vector<int> idx;
};
struct Poly {
    std::vector<F2> p, vel;
    bool active = false;
};
struct Frame2 {
    double dur = 0;
    std::vector<Poly> shapes;
};

template <class T>
T* dalloc(size_t n) {
    T* p = nullptr;
    CMCX_HIP(hipMalloc((void**)&p, sizeof(T) * std::max<size_t>(n, 1)));
    return p;
}

unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

struct Geometry {
    int kind = 0;  // 3 = triangle mesh, 2 = extruded polylines
    double h[3] = {0, 0, 0}, base_t = 1.0, depth = 0, depth_var = 0;
    int nx = 0, ny = 0, nz = 0, active_nz = 0;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    std::vector<Frame3> f3;
    std::vector<Frame2> f2;
    std::vector<double> dur;
    int cx = 0, cy = 0;  // 2D grid
    // device node arrays (kept between builds: walls of a mesh keep their data)
    uint8_t *type = nullptr, *bcv = nullptr, *bct = nullptr;
    float* fv[4] = {nullptr, nullptr, nullptr, nullptr};
    int* d_changed = nullptr;

    ~Geometry() {
        hipFree(type);
        hipFree(bcv);
        hipFree(bct);
        for (float* p : fv) hipFree(p);
        hipFree(d_changed);
    }

    int64_t N() const { return (int64_t)nx * ny * nz; }

    void alloc_nodes() {
        const int64_t n = N();
        type = dalloc<uint8_t>(n);
        bcv = dalloc<uint8_t>(n);
        bct = dalloc<uint8_t>(n);
        for (float*& p : fv) p = dalloc<float>(n);
        // Init: NODE_OUT, v = 0, T = 0 (BC kinds zero: BC_NOSLIP)
        hipLaunchKernelGGL(fill_u8_kernel, dim3(nblk(n)), dim3(256), 0, 0, type, n, (uint8_t)NODE_OUT);
        CMCX_HIP(hipMemset(bcv, 0, n));
        CMCX_HIP(hipMemset(bct, 0, n));
        for (float* p : fv) CMCX_HIP(hipMemset(p, 0, sizeof(float) * n));
        CMCX_HIP(hipDeviceSynchronize());
    }

    // frame and blend weight at `time` (Grid3D::Prepare3D_Shape, Grid2D::Prepare)
    void frame_at(double time, int& fr, double& sub) const {
        std::vector<double> a(dur.size() + 1, 0.0);
        for (size_t i = 1; i <= dur.size(); i++) a[i] = a[i - 1] + dur[i - 1];
        const double r = std::fmod(time, a[dur.size()]);
        fr = 0;
        for (size_t i = 1; i < dur.size(); i++)
            if (a[i] < r) fr = (int)i;
        sub = (r - a[fr]) / (a[fr + 1] - a[fr]);
    }

    // iterated flood fill from node 0 over a grid of nx x ny x nz (nz = 1: 2D)
    void flood(uint8_t* t, Box b) {
        if (!d_changed) d_changed = dalloc<int>(1);
        hipLaunchKernelGGL(fill_u8_kernel, dim3(1), dim3(1), 0, 0, t, (int64_t)1, (uint8_t)NODE_OUT);
        const int64_t n = (int64_t)b.nx * b.ny * b.nz;
        const unsigned g = (unsigned)std::min<int64_t>(nblk(n), 2048);
        for (int round = 0; round < (int)(n + 1); round += 4) {
            CMCX_HIP(hipMemset(d_changed, 0, sizeof(int)));
            for (int r = 0; r < 4; r++) hipLaunchKernelGGL(flood_kernel, dim3(g), dim3(256), 0, 0, t, b, d_changed);
            int h_changed = 0;
            CMCX_HIP(hipMemcpy(&h_changed, d_changed, sizeof(int), hipMemcpyDeviceToHost));
            if (!h_changed) break;
        }
    }

    void build3(double time) {
        int fr;
        double sub;
        frame_at(time, fr, sub);
        const Frame3& a = f3[fr % f3.size()];
        const Frame3& b = f3[(fr + 1) % f3.size()];
        const float ss = (float)sub, is = 1 - ss;
        std::vector<F3> v(a.v.size());
        for (size_t k = 0; k < v.size(); k++)
            v[k] = F3{a.v[k].x * is + b.v[k].x * ss, a.v[k].y * is + b.v[k].y * ss,
                      a.v[k].z * is + b.v[k].z * ss};
        const int ntri = (int)(a.idx.size() / 3);
        F3* dv = dalloc<F3>(v.size());
        int* di = dalloc<int>(a.idx.size());
        CMCX_HIP(hipMemcpy(dv, v.data(), sizeof(F3) * v.size(), hipMemcpyHostToDevice));
        CMCX_HIP(hipMemcpy(di, a.idx.data(), sizeof(int) * a.idx.size(), hipMemcpyHostToDevice));
        const Box box{nx, ny, nz};
        hipLaunchKernelGGL(fill_u8_kernel, dim3(nblk(N())), dim3(256), 0, 0, type, N(), (uint8_t)NODE_IN);
        if (ntri) hipLaunchKernelGGL(raster_tri_kernel, dim3(nblk(ntri, 64)), dim3(64), 0, 0, dv, di, ntri, box, type);
        CMCX_LAUNCH_CHECK();
        flood(type, box);
        hipLaunchKernelGGL(mesh_data_kernel, dim3(nblk(N())), dim3(256), 0, 0, type, N(), bcv, bct, fv[0],
                           fv[1], fv[2], fv[3], (float)base_t);
        CMCX_LAUNCH_CHECK();
        CMCX_HIP(hipDeviceSynchronize());
        hipFree(dv);
        hipFree(di);
    }

    // Grid2D::Prepare(time) into device cell arrays (cell type, v, T; the BC kind
    // of every cell is no-slip); the caller frees them.
    struct Cells2 {
        uint8_t* cell = nullptr;
        float *vx = nullptr, *vy = nullptr, *T = nullptr;
        ~Cells2() { hipFree(cell); hipFree(vx); hipFree(vy); hipFree(T); }
    };

    void build2(double time, Cells2* keep = nullptr) {
        int fr;
        double sub;
        frame_at(time, fr, sub);
        const Frame2& a = f2[fr % f2.size()];
        const Frame2& b = f2[(fr + 1) % f2.size()];
        const double isd = 1 - sub;
        const float is = (float)isd, ss = (float)sub;
        // segments in the reference's order: the moving shapes' (valves), then the walls
        std::vector<Seg2> segs;
        for (int pass = 0; pass < 2; pass++)
            for (size_t s = 0; s < a.shapes.size(); s++) {
                const Poly& pa = a.shapes[s];
                const Poly& pb = b.shapes[s];
                if (pa.active != (pass == 0)) continue;
                std::vector<F2> p(pa.p.size()), vel(pa.p.size());
                for (size_t k = 0; k < pa.p.size(); k++) {
                    p[k] = F2{pa.p[k].x * is + pb.p[k].x * ss, pa.p[k].y * is + pb.p[k].y * ss};
                    vel[k] = F2{pa.vel[k].x * is + pb.vel[k].x * ss, pa.vel[k].y * is + pb.vel[k].y * ss};
                }
                for (size_t k = 0; k + 1 < p.size(); k++)
                    segs.push_back(Seg2{p[k], p[k + 1], vel[k], vel[k + 1], pass == 0 ? NODE_VALVE : NODE_BOUND});
            }
        const int64_t nc = (int64_t)cx * cy;
        Cells2 own;
        Cells2& C = keep ? *keep : own;
        C.cell = dalloc<uint8_t>(nc);
        C.vx = dalloc<float>(nc);
        C.vy = dalloc<float>(nc);
        C.T = dalloc<float>(nc);
        uint8_t* cell = C.cell;
        float* cvx = C.vx;
        float* cvy = C.vy;
        float* cT = C.T;
        unsigned long long* key = dalloc<unsigned long long>(nc);
        Seg2* ds = dalloc<Seg2>(std::max<size_t>(segs.size(), 1));
        try {
            CMCX_HIP(hipMemcpy(ds, segs.data(), sizeof(Seg2) * segs.size(), hipMemcpyHostToDevice));
            CMCX_HIP(hipMemset(cvx, 0, sizeof(float) * nc));
            CMCX_HIP(hipMemset(cvy, 0, sizeof(float) * nc));
            CMCX_HIP(hipMemset(cT, 0, sizeof(float) * nc));
            CMCX_HIP(hipMemset(key, 0, sizeof(unsigned long long) * nc));
            hipLaunchKernelGGL(fill_u8_kernel, dim3(nblk(nc)), dim3(256), 0, 0, cell, nc, (uint8_t)NODE_IN);
            const int ns = (int)segs.size();
            if (ns) {
                for (int mode = 0; mode < 2; mode++)
                    hipLaunchKernelGGL(raster_seg2_kernel, dim3(nblk(ns, 64)), dim3(64), 0, 0, ds, ns, cx, cy,
                                       mode, key, cell, cvx, cvy, cT, (float)base_t);
            }
            CMCX_LAUNCH_CHECK();
            flood(cell, Box{cx, cy, 1});
            hipLaunchKernelGGL(cell_data_kernel, dim3(nblk(nc)), dim3(256), 0, 0, cell, nc, cvx, cvy, cT,
                               (float)base_t);
            if (nz > 0)
                hipLaunchKernelGGL(extrude_kernel, dim3(nblk(nc, 64)), dim3(64), 0, 0, cell, cvx, cvy, cT,
                                   Box{nx, ny, nz}, active_nz, depth_var, (float)base_t, type, bcv, bct,
                                   fv[0], fv[1], fv[2], fv[3]);
            CMCX_LAUNCH_CHECK();
            CMCX_HIP(hipDeviceSynchronize());
        } catch (...) {
            hipFree(key); hipFree(ds);
            throw;
        }
        hipFree(key); hipFree(ds);
    }
};

namespace {

bool parse_mesh(const char* path, Geometry& G, std::string& err) {
    Text tx;
    if (!tx.load(path)) { err = "cannot open file"; return false; }
    int nf = 0;
    if (!tx.integer(nf) || nf < 1) { err = "bad frame count"; return false; }
    G.f3.resize(nf);
    G.dur.assign(nf, 1.0 / 75);  // 75 frames per second (Grid3D.cpp:419)
    for (Frame3& f : G.f3) {
        int nv = 0, ni = 0;
        if (!tx.integer(nv) || nv < 0) { err = "bad vertex count"; return false; }
        f.v.resize(nv);
        for (int k = 0; k < nv; k++) {
            float c[6];
            for (float& x : c) {
                std::string w = tx.word();
                if (w.empty()) { err = "truncated vertex list"; return false; }
                x = to_float(w);
            }
            // position (file units to metres); the vertex velocity is read and, as
            // in the reference, not used by the build
            f.v[k] = F3{c[0] * kScale, c[1] * kScale, c[2] * kScale};
        }
        if (!tx.integer(ni) || ni < 0) { err = "bad triangle count"; return false; }
        f.idx.resize(3 * (size_t)ni);
        for (int& x : f.idx)
            if (!tx.integer(x) || x < 0 || x >= nv) { err = "bad vertex index"; return false; }
    }
    // frames are blended vertex by vertex: every frame lists the same vertices
    for (const Frame3& f : G.f3)
        if (f.v.size() != G.f3[0].v.size() || f.idx.size() != G.f3[0].idx.size()) {
            err = "frames differ in vertex or triangle count";
            return false;
        }
    // BBox3D::Build with 2 % padding, then Init
    float mn[3] = {1e10f, 1e10f, 1e10f}, mx[3] = {-1e10f, -1e10f, -1e10f};
    for (const Frame3& f : G.f3)
        for (const F3& p : f.v) {
            const float c[3] = {p.x, p.y, p.z};
            for (int a = 0; a < 3; a++) {
                if (c[a] < mn[a]) mn[a] = c[a];
                if (c[a] > mx[a]) mx[a] = c[a];
            }
        }
    for (int a = 0; a < 3; a++) {
        const float w = mx[a] - mn[a];
        mn[a] -= w * (float)kPad;
        mx[a] += w * (float)kPad;
        G.bmin[a] = mn[a];
        G.bmax[a] = mx[a];
    }
    G.nx = (int)std::ceil((mx[0] - mn[0]) / G.h[0]) + 1;
    G.ny = (int)std::ceil((mx[1] - mn[1]) / G.h[1]) + 1;
    G.nz = (int)std::ceil((mx[2] - mn[2]) / G.h[2]) + 1;
    for (Frame3& f : G.f3)
        for (F3& p : f.v) {
            p.x = (p.x - mn[0]) / (float)G.h[0];
            p.y = (p.y - mn[1]) / (float)G.h[1];
            p.z = (p.z - mn[2]) / (float)G.h[2];
        }
    return true;
}

bool parse_polylines(const char* path, Geometry& G, std::string& err) {
    Text tx;
    if (!tx.load(path)) { err = "cannot open file"; return false; }
    int nf = 0;
    if (!tx.integer(nf) || nf < 1) { err = "bad frame count"; return false; }
    G.f2.resize(nf);
    for (Frame2& f : G.f2) {
        std::string w = tx.word();
        if (w.empty()) { err = "missing frame duration"; return false; }
        f.dur = (double)strtof(w.c_str(), nullptr);  // read as "%f" into a float
        int ns = 0;
        if (!tx.integer(ns) || ns < 0) { err = "bad shape count"; return false; }
        f.shapes.resize(ns);
        for (Poly& s : f.shapes) {
            int np = 0;
            if (!tx.integer(np) || np < 0) { err = "bad point count"; return false; }
            s.p.resize(np);
            s.vel.resize(np);
            for (F2& p : s.p) {
                const F2 q = point2(tx);
                p = F2{q.x * kScale, q.y * kScale};
            }
            const std::string kind = tx.word();
            F2 v{0, 0};
            s.active = !kind.empty() && kind[0] == 'M';  // "Motion <vx vy>" or "Passive"
            if (s.active) v = point2(tx);
            for (F2& x : s.vel) x = F2{v.x * kScale, v.y * kScale};
        }
    }
    // frames are blended point by point: every frame lists the same polylines
    for (size_t k = 1; k < G.f2.size(); k++) {
        if (G.f2[k].shapes.size() != G.f2[0].shapes.size()) { err = "frames differ in shapes"; return false; }
        for (size_t i = 0; i < G.f2[0].shapes.size(); i++)
            if (G.f2[k].shapes[i].p.size() != G.f2[0].shapes[i].p.size()) {
                err = "frames differ in polyline points";
                return false;
            }
    }
    // ComputeBorderVelocities (Grid2D.cpp:358-379)
    for (int fr = 0; fr < nf; fr++) {
        Frame2& a = G.f2[fr];
        Frame2& b = G.f2[(fr + 1) % nf];
        const float m = (float)(1 / a.dur);
        for (size_t i = 0; i < a.shapes.size(); i++)
            for (size_t k = 0; k < a.shapes[i].p.size(); k++) {
                const F2 pa = a.shapes[i].p[k], pb = b.shapes[i].p[k];
                if (!a.shapes[i].active) {
                    b.shapes[i].vel[k] = F2{(pb.x - pa.x) * m, (pb.y - pa.y) * m};
                } else {
                    b.shapes[i].vel[k].x += (pa.x - pb.x) * m;
                    b.shapes[i].vel[k].y += (pa.y - pb.y) * m;
                }
            }
    }
    G.dur.clear();
    for (const Frame2& f : G.f2) G.dur.push_back(f.dur);
    // BBox2D::Build and Grid2D::Init
    float mn[2] = {1e10f, 1e10f}, mx[2] = {-1e10f, -1e10f};
    for (const Frame2& f : G.f2)
        for (const Poly& s : f.shapes)
            for (const F2& p : s.p) {
                if (p.x < mn[0]) mn[0] = p.x;
                if (p.y < mn[1]) mn[1] = p.y;
                if (p.x > mx[0]) mx[0] = p.x;
                if (p.y > mx[1]) mx[1] = p.y;
            }
    for (int a = 0; a < 2; a++) {
        const float w = mx[a] - mn[a];
        mn[a] -= w * (float)kPad;
        mx[a] += w * (float)kPad;
    }
    // the output box of an extruded shape: the 2D box, depth below z = 0
    G.bmin[0] = mn[0]; G.bmin[1] = mn[1]; G.bmin[2] = -(float)G.depth;
    G.bmax[0] = mx[0]; G.bmax[1] = mx[1]; G.bmax[2] = 0.0f;
    G.cx = (int)std::ceil((mx[0] - mn[0]) / G.h[0]) + 1;
    G.cy = (int)std::ceil((mx[1] - mn[1]) / G.h[1]) + 1;
    for (Frame2& f : G.f2)
        for (Poly& s : f.shapes)
            for (F2& p : s.p) {
                p.x = (p.x - mn[0]) / (float)G.h[0];
                p.y = (p.y - mn[1]) / (float)G.h[1];
            }
    return true;
}

}  // namespace
}  // namespace cmcx

using namespace cmcx;

struct cmcx_geometry {
    Geometry g;
};

extern "C" {

int cmcx_geometry_load(int kind, const char* path, const double* spacing, double base_t, double depth,
                       double depth_var, int align, cmcx_geometry** out) {
    if (!path || !spacing || !out) return CMCX_ERR_ARG;
    *out = nullptr;
    if (kind != CMCX_SHAPE3D && kind != CMCX_SHAPE2D) return CMCX_ERR_ARG;
    for (int a = 0; a < 3; a++)
        if (!(spacing[a] > 0)) {
            set_error("cmcx_geometry_load: grid spacings must be > 0");
            return CMCX_ERR_ARG;
        }
    if (kind == CMCX_SHAPE2D && depth < 0) {
        set_error("cmcx_geometry_load: depth must be >= 0 (0: the 2D grid only)");
        return CMCX_ERR_ARG;
    }
    auto* h = new cmcx_geometry;
    Geometry& G = h->g;
    G.kind = kind == CMCX_SHAPE3D ? 3 : 2;
    for (int a = 0; a < 3; a++) G.h[a] = spacing[a];
    G.base_t = base_t;
    G.depth = depth;
    G.depth_var = depth_var;
    std::string err;
    const bool ok = G.kind == 3 ? parse_mesh(path, G, err) : parse_polylines(path, G, err);
    if (!ok) {
        delete h;
        set_error(std::string("cmcx_geometry_load: ") + path + ": " + err);
        return CMCX_ERR_ARG;
    }
    if (G.kind == 3) {
        if (align) { G.nx = align32(G.nx); G.ny = align32(G.ny); G.nz = align32(G.nz); }
    } else {
        if (align) { G.cx = align32(G.cx); G.cy = align32(G.cy); }
        G.nx = G.cx;
        G.ny = G.cy;
        if (depth > 0) {
            G.active_nz = (int)std::ceil(depth / spacing[2]) + 1;
            G.nz = align ? align32(G.active_nz) : G.active_nz;
        }
    }
    const bool only2d = G.kind == 2 && G.nz == 0;
    if (G.nx < 1 || G.ny < 1 || (G.nz < 1 && !only2d) || G.N() > ((int64_t)1 << 31) - 1 ||
        (int64_t)G.nx * G.ny > ((int64_t)1 << 31) - 1) {
        delete h;
        set_error("cmcx_geometry_load: grid empty or too large");
        return CMCX_ERR_TOO_LARGE;
    }
    try {
        if (!only2d) G.alloc_nodes();
    } catch (const HipError& e) {
        delete h;
        set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
        return CMCX_ERR_HIP;
    }
    *out = h;
    return CMCX_OK;
}

int cmcx_geometry_info(cmcx_geometry* h, int32_t* dims, double* bbox_min, double* bbox_max, int* nframes,
                       double* cycle) {
    if (!h) return CMCX_ERR_ARG;
    const Geometry& G = h->g;
    if (dims) { dims[0] = G.nx; dims[1] = G.ny; dims[2] = G.nz; }
    for (int a = 0; a < 3; a++) {
        if (bbox_min) bbox_min[a] = G.bmin[a];
        if (bbox_max) bbox_max[a] = G.bmax[a];
    }
    if (nframes) *nframes = (int)G.dur.size();
    if (cycle) {
        double c = 0;
        for (double d : G.dur) c += d;
        *cycle = c;
    }
    return CMCX_OK;
}

int cmcx_geometry_build(cmcx_geometry* h, double time, uint8_t* type, uint8_t* bc_vel, uint8_t* bc_temp,
                        double* vx, double* vy, double* vz, double* T) {
    if (!h) return CMCX_ERR_ARG;
    Geometry& G = h->g;
    if (G.nz == 0) {
        set_error("cmcx_geometry_build: loaded with depth 0 (the 2D grid only): use cmcx_geometry_build2d");
        return CMCX_ERR_STATE;
    }
    try {
        if (G.kind == 3) G.build3(time);
        else G.build2(time);
        const int64_t n = G.N();
        uint8_t* u8[3] = {type, bc_vel, bc_temp};
        const uint8_t* d8[3] = {G.type, G.bcv, G.bct};
        for (int a = 0; a < 3; a++)
            if (u8[a]) CMCX_HIP(hipMemcpy(u8[a], d8[a], n, hipMemcpyDeviceToHost));
        double* dv[4] = {vx, vy, vz, T};
        std::vector<float> tmp;
        for (int f = 0; f < 4; f++)
            if (dv[f]) {
                tmp.resize(n);
                CMCX_HIP(hipMemcpy(tmp.data(), G.fv[f], sizeof(float) * n, hipMemcpyDeviceToHost));
                for (int64_t q = 0; q < n; q++) dv[f][q] = (double)tmp[q];
            }
    } catch (const HipError& e) {
        set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
        return CMCX_ERR_HIP;
    }
    return CMCX_OK;
}

int cmcx_geometry_build2d(cmcx_geometry* h, double time, uint8_t* type, uint8_t* bc, double* vx, double* vy,
                          double* T) {
    if (!h) return CMCX_ERR_ARG;
    Geometry& G = h->g;
    if (G.kind != 2) {
        set_error("cmcx_geometry_build2d: a polyline (CMCX_SHAPE2D) input is needed");
        return CMCX_ERR_ARG;
    }
    try {
        Geometry::Cells2 C;
        G.build2(time, &C);
        const int64_t n = (int64_t)G.cx * G.cy;
        if (type) CMCX_HIP(hipMemcpy(type, C.cell, n, hipMemcpyDeviceToHost));
        if (bc) std::memset(bc, 0, n);  // BC_NOSLIP (Grid2D::RasterLine with bc_noslip)
        const float* src[3] = {C.vx, C.vy, C.T};
        double* dst[3] = {vx, vy, T};
        std::vector<float> tmp;
        for (int f = 0; f < 3; f++)
            if (dst[f]) {
                tmp.resize(n);
                CMCX_HIP(hipMemcpy(tmp.data(), src[f], sizeof(float) * n, hipMemcpyDeviceToHost));
                for (int64_t q = 0; q < n; q++) dst[f][q] = (double)tmp[q];
            }
    } catch (const HipError& e) {
        set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
        return CMCX_ERR_HIP;
    }
    return CMCX_OK;
}

int cmcx_geometry_destroy(cmcx_geometry* h) {
    if (!h) return CMCX_ERR_ARG;
    delete h;
    return CMCX_OK;
}

}  // extern "C"
