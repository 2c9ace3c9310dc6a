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
// Algorithms.h:21-38 in the reference's operation order (c', d' in a scratch
// laid out (row, field, line) so neighbouring threads touch neighbouring
// words); copies, merges and the divergence sum are elementwise.  The grids of
// this path are small (256 x 128 in the reference configs), so the step is
// launch-bound: each Picard iteration is 13 launches and one host read of the
// error (the loop condition needs it).
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

// One line system per thread (blockIdx.y = field): BuildMatrix
// (AdiSolver2D.cpp:118-178), ApplyBC0/1 (:74-116), SolveTridiagonal, UpdateSegment.
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
    const T dt = (T)prm.dt, hx = (T)prm.dx, hy = (T)prm.dy;
    const T v_T = (T)prm.v_T, t_phi = (T)prm.t_phi;
    const T v_vis = (T)prm.v_vis, t_vis = (T)prm.t_vis;
    const T vis = DIRX ? (var == 2 ? t_vis / (hx * hx) : v_vis / (hx * hx))
                       : (var == 2 ? t_vis / (hy * hy) : v_vis / (hy * hy));
    const T h = DIRX ? hx : hy;
    const T* U = tl.f[0];
    const T* V = tl.f[1];
    const T* TT = tl.f[2];
    const int64_t sxs = ny;  // stride of i
    auto idx = [&](int t) { return ((int64_t)t * 3 + var) * nsys + s; };  // (row, field, line)
    // row 0: ApplyBC0
    T a, b, c, d;
    {
        const int64_t c0 = cell(0);
        if (bc[c0] == BC_NOSLIP) {
            b = 1; c = 0; d = g.f[var][c0];
        } else {
            b = 1; c = -1; d = 0;
        }
    }
    // forward elimination, rows built on the fly (SolveTridiagonal with c[n-1] = 0)
    T cp = c / b, dp = d / b;
    sc[idx(0)] = cp;
    sd[idx(0)] = dp;
    for (int t = 1; t < n; t++) {
        const int64_t ct = cell(t);
        if (t < n - 1) {
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

__global__ void k2_div_finish(const double* __restrict__ part, int nb, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0, c = 0;
    for (int b = 0; b < nb; b++) {
        s += part[2 * b];
        c += part[2 * b + 1];
    }
    out[0] = s;
    out[1] = c;
}

}  // namespace

struct ISolver2 {
    virtual ~ISolver2() {}
    virtual int update_boundaries() = 0;
    virtual int step(double dt, int G, int L, double* err, int* iters) = 0;
    virtual int get_fields(void* u, void* v, void* p, void* T) = 0;
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
    T* scratch = nullptr;
    int64_t sc_half = 0;
    double* d_part = nullptr;
    double* d_err = nullptr;

    Solver2D(const cmcx2d_domain* d, const cmcx_params* p, hipStream_t stream) {
        st = stream;
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
        build_systems(d->type);
        CMCX_HIP(hipMalloc((void**)&d_part, sizeof(double) * 2 * kDiv2Blocks));
        CMCX_HIP(hipMalloc((void**)&d_err, sizeof(double) * 2));
        CMCX_HIP(hipStreamSynchronize(st));
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

    void sweep(bool dirx, L3<T> lc, L3<T> lt, L3<T> out) {
        const int n = dirx ? nrows : ncols;
        if (!n) return;
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

    // AdiSolver2D::SolveDirection
    void direction(bool dirx, int L, L3<T> lc, L3<T> lt, L3<T> out) {
        if (!(dirx ? nrows : ncols)) return;  // the reference reads list[0] unconditionally
        copy(lt, tl, kAll, 0);
        for (int it = 0; it < L; it++) {
            sweep(dirx, lc, tl, nl);
            copy(nl, tl, kIn, it == 0 ? 0 : 1);
        }
        copy(tl, lt, kIn, 0);
        copy(nl, out, kIn, 0);
    }

    double div_error(L3<T> Lr) {
        hipLaunchKernelGGL((k2_div<T>), dim3(kDiv2Blocks), dim3(kBS2), 0, st, nx, ny, d_type, Lr,
                           (T)prm.dx, (T)prm.dy, d_part);
        CMCX_LAUNCH_CHECK();
        hipLaunchKernelGGL(k2_div_finish, dim3(1), dim3(64), 0, st, d_part, kDiv2Blocks, d_err);
        CMCX_LAUNCH_CHECK();
        double h[2];
        CMCX_HIP(hipMemcpyAsync(h, d_err, sizeof(h), hipMemcpyDeviceToHost, st));
        CMCX_HIP(hipStreamSynchronize(st));
        return h[0] / h[1];
    }

    int update_boundaries() override {
        hipLaunchKernelGGL((k2_update_boundaries<T>), dim3(blocks2(N)), dim3(kBS2), 0, st, nx, ny,
                           d_type, g, cur, next);
        CMCX_LAUNCH_CHECK();
        return CMCX_OK;
    }

    // AdiSolver2D::TimeStep (AdiSolver2D.cpp:279-323)
    int step(double dt, int G, int L, double* err_out, int* iters) override {
        if (G < 0 || L < 1) return CMCX_ERR_ARG;
        prm.dt = dt;
        copy(cur, next, kAll, 0);
        copy(cur, half, kAll, 0);
        copy(cur, temp, kAll, 0);
        double err = div_error(next);
        int it;
        for (it = 0; it < G || err > 0.1; it++) {
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
    if (!d || !p || !out) return CMCX_ERR_ARG;
    *out = nullptr;
    if (dtype != CMCX_F32 && dtype != CMCX_F64) return CMCX_ERR_ARG;
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
            if (dtype == CMCX_F32) s->impl.reset(new cmcx::Solver2D<float>(d, p, (hipStream_t)stream));
            else s->impl.reset(new cmcx::Solver2D<double>(d, p, (hipStream_t)stream));
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

int cmcx2d_destroy(cmcx2d_solver* s) {
    if (!s) return CMCX_ERR_ARG;
    CMCX2_GUARD({
        delete s;
        return CMCX_OK;
    })
}

}  // extern "C"
