// ORACLE -- test infrastructure only.
//
// CPU restatement of the reference 2D ADI solver (src/FluidSolver2D, CPU only
// in the reference).  Only tests/ and bench.py's cpu_baseline leg may load it.
//
// Follows:
//   create_segments   src/FluidSolver2D/AdiSolver2D.cpp:228-277
//   apply_bc0/1       src/FluidSolver2D/AdiSolver2D.cpp:74-116 (free = zero gradient)
//   build_matrix      src/FluidSolver2D/AdiSolver2D.cpp:118-178 (1/dt, not 3/dt)
//   solve_direction   src/FluidSolver2D/AdiSolver2D.cpp:203-226
//   time_step         src/FluidSolver2D/AdiSolver2D.cpp:279-323
//   update_boundaries src/FluidSolver2D/Solver2D.cpp:48-63
//   clear_outer       src/FluidSolver2D/Solver2D.cpp:75-85
//   copy/merge/div    src/FluidSolver2D/TimeLayer2D.h:88-172
//   thomas            src/Common/Algorithms.h:21-38
//
// The Stable solver (projection method, src/FluidSolver2D/StableSolver2D.cpp):
//   stable_init       StableSolver2D.cpp:21-53 (next and temp copied from the state)
//   prepare_indices   StableSolver2D.cpp:167-188
//   solve_uv          StableSolver2D.cpp:55-81 (explicit advection-diffusion)
//   project           StableSolver2D.cpp:83-165 (Poisson relaxation, gradient subtraction)
//   stable_step       StableSolver2D.cpp:190-234
//   grad              ScalarField2D.h:29-36
// project's relaxation updates q in place, boundary list then inner list, each
// in (i, j) order: a Gauss-Seidel sweep.  Its stopping test folds the cells'
// relative changes with std::max(cur, err) = (cur < err) ? err : cur, so a 0/0
// change (NaN) resets the running maximum and a NaN on the last cell ends the
// loop.  Grid2D::GetType and the q fields have no bounds check
// (Grid2D.cpp:50-53): a boundary cell on the grid's edge reads outside the
// arrays.  Here the grid is taken as surrounded by NODE_OUT cells with q = 0.
//
// Layers the reference allocates with new[] and never fully writes (its copy
// and merge loops stop at dimx-2 / dimy-2, TimeLayer2D.h:106-150) are zero
// here: every layer, including the per-direction temp_local, starts zeroed.
// The reference reads those cells (neighbours of the last row/column) from
// fresh heap memory, which glibc hands out zeroed on first use.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc2 {

enum { NODE_IN = 0, NODE_OUT = 1, NODE_BOUND = 2, NODE_VALVE = 3 };
enum { BC_NOSLIP = 0, BC_FREE = 1 };
enum { DIR_X = 0, DIR_Y = 1 };
enum { VAR_U = 0, VAR_V = 1, VAR_T = 2 };

struct Seg2 { int px, py, ex, ey, size, dir; };

template <class T>
struct Layer2 {
    int nx, ny;
    T dx, dy;
    std::vector<T> u, v, t;
    void init(int x, int y, T hx, T hy) {
        nx = x; ny = y; dx = hx; dy = hy;
        u.assign((size_t)x * y, (T)0); v.assign((size_t)x * y, (T)0); t.assign((size_t)x * y, (T)0);
    }
    T& U(int i, int j) { return u[(size_t)i * ny + j]; }
    T& V(int i, int j) { return v[(size_t)i * ny + j]; }
    T& Tt(int i, int j) { return t[(size_t)i * ny + j]; }
    T& F(int var, int i, int j) { return var == VAR_U ? U(i, j) : var == VAR_V ? V(i, j) : Tt(i, j); }
    T Ux(int i, int j) { return (U(i + 1, j) - U(i - 1, j)) / (2 * dx); }
    T Uy(int i, int j) { return (U(i, j + 1) - U(i, j - 1)) / (2 * dy); }
    T Vx(int i, int j) { return (V(i + 1, j) - V(i - 1, j)) / (2 * dx); }
    T Vy(int i, int j) { return (V(i, j + 1) - V(i, j - 1)) / (2 * dy); }
    T Uxx(int i, int j) { return (U(i + 1, j) - 2 * U(i, j) + U(i - 1, j)) / (dx * dx); }
    T Uyy(int i, int j) { return (U(i, j + 1) - 2 * U(i, j) + U(i, j - 1)) / (dy * dy); }
    T Vxx(int i, int j) { return (V(i + 1, j) - 2 * V(i, j) + V(i - 1, j)) / (dx * dx); }
    T Vyy(int i, int j) { return (V(i, j + 1) - 2 * V(i, j) + V(i, j - 1)) / (dy * dy); }
    T Tx(int i, int j) { return (Tt(i + 1, j) - Tt(i - 1, j)) / (2 * dx); }
    T Ty(int i, int j) { return (Tt(i, j + 1) - Tt(i, j - 1)) / (2 * dy); }
    T DissFuncX(int i, int j) {
        T ux = Ux(i, j), vx = Vx(i, j), uy = Uy(i, j);
        return 2 * ux * ux + vx * vx + uy * vx;
    }
    T DissFuncY(int i, int j) {
        T vx = Vx(i, j), uy = Uy(i, j), vy = Vy(i, j);
        return uy * uy + 2 * vy * vy + vx * uy;
    }
};

template <class T>
struct Solver2 {
    int nx, ny;
    double gdx, gdy, startT;
    std::vector<uint8_t> type, bc;
    std::vector<T> gvx, gvy, gT;
    T v_T, v_vis, t_vis, t_phi;
    Layer2<T> cur, next, half, temp, next_local;
    std::vector<Seg2> listX, listY;  // listX: one Y-segment per column i; listY: one X-segment per row j

    uint8_t GetType(int i, int j) { return type[(size_t)i * ny + j]; }
    size_t id(int i, int j) { return (size_t)i * ny + j; }

    // copy of one var family over cells of one type, i < nx-1, j < ny-1 (TimeLayer2D.h:104-172)
    void copy_type(Layer2<T>& src, Layer2<T>& dst, int t) {
        for (int i = 0; i < nx - 1; i++)
            for (int j = 0; j < ny - 1; j++)
                if (GetType(i, j) == t) {
                    dst.U(i, j) = src.U(i, j); dst.V(i, j) = src.V(i, j); dst.Tt(i, j) = src.Tt(i, j);
                }
    }
    void copy_all(Layer2<T>& src, Layer2<T>& dst) {
        for (int t : {NODE_IN, NODE_OUT, NODE_BOUND, NODE_VALVE}) copy_type(src, dst, t);
    }
    void merge_type(Layer2<T>& src, Layer2<T>& dst, int t) {
        for (int i = 0; i < nx - 1; i++)
            for (int j = 0; j < ny - 1; j++)
                if (GetType(i, j) == t) {
                    dst.U(i, j) = (dst.U(i, j) + src.U(i, j)) / 2;
                    dst.V(i, j) = (dst.V(i, j) + src.V(i, j)) / 2;
                    dst.Tt(i, j) = (dst.Tt(i, j) + src.Tt(i, j)) / 2;
                }
    }

    // TimeLayer2D::EvalDivError (TimeLayer2D.h:88-102): FTYPE accumulation.
    double div_error(Layer2<T>& L) {
        T err = 0.0;
        int count = 0;
        for (int i = 0; i < nx - 1; i++)
            for (int j = 0; j < ny - 1; j++)
                if (GetType(i, j) == NODE_IN && GetType(i + 1, j) == NODE_IN &&
                    GetType(i, j + 1) == NODE_IN && GetType(i + 1, j + 1) == NODE_IN) {
                    T tx = L.dy * (L.U(i + 1, j) - L.U(i, j)) + (L.U(i + 1, j + 1) - L.U(i, j + 1)) / 2;
                    T ty = L.dx * (L.V(i, j + 1) - L.V(i, j)) + (L.V(i + 1, j + 1) - L.V(i + 1, j)) / 2;
                    err += std::fabs(tx + ty);
                    count++;
                }
        return (double)(err / count);
    }

    void init(int x, int y, double hx, double hy, double sT, const uint8_t* t, const uint8_t* b,
              const double* vx, const double* vy, const double* TT) {
        nx = x; ny = y; gdx = hx; gdy = hy; startT = sT;
        size_t N = (size_t)x * y;
        type.assign(t, t + N); bc.assign(b, b + N);
        gvx.resize(N); gvy.resize(N); gT.resize(N);
        for (size_t i = 0; i < N; i++) { gvx[i] = (T)vx[i]; gvy[i] = (T)vy[i]; gT[i] = (T)TT[i]; }
        for (auto* L : {&cur, &next, &half, &temp, &next_local}) L->init(x, y, (T)hx, (T)hy);
        // AdiSolver2D::Init (AdiSolver2D.cpp:37-49): all cells from the grid
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) {
                cur.U(i, j) = gvx[id(i, j)]; cur.V(i, j) = gvy[id(i, j)]; cur.Tt(i, j) = gT[id(i, j)];
            }
    }

    void create_segments() {
        listX.clear();
        for (int i = 0; i < nx; i++) {
            Seg2 s; s.px = i; s.dir = DIR_Y;
            int j = 0;
            while (j < ny && GetType(i, j) == NODE_OUT) j++;
            while (j + 1 < ny && GetType(i, j + 1) != NODE_IN) j++;
            if (j + 1 >= ny) continue;
            s.py = j;
            j = ny - 1;
            while (j >= 0 && GetType(i, j) == NODE_OUT) j--;
            while (j - 1 >= 0 && GetType(i, j - 1) != NODE_IN) j--;
            s.size = j - s.py + 1; s.ex = i; s.ey = j;
            listX.push_back(s);
        }
        listY.clear();
        for (int j = 0; j < ny; j++) {
            Seg2 s; s.py = j; s.dir = DIR_X;
            int i = 0;
            while (i < nx && GetType(i, j) == NODE_OUT) i++;
            while (i + 1 < nx && GetType(i + 1, j) != NODE_IN) i++;
            if (i + 1 >= nx) continue;
            s.px = i;
            i = nx - 1;
            while (i >= 0 && GetType(i, j) == NODE_OUT) i--;
            while (i - 1 >= 0 && GetType(i - 1, j) != NODE_IN) i--;
            s.size = i - s.px + 1; s.ex = i; s.ey = j;
            listY.push_back(s);
        }
    }

    static void thomas(T* a, T* b, T* c, T* d, T* x, int num) {
        c[num - 1] = 0.0;
        c[0] = c[0] / b[0];
        d[0] = d[0] / b[0];
        for (int i = 1; i < num; i++) {
            c[i] = c[i] / (b[i] - a[i] * c[i - 1]);
            d[i] = (d[i] - d[i - 1] * a[i]) / (b[i] - a[i] * c[i - 1]);
        }
        x[num - 1] = d[num - 1];
        for (int i = num - 2; i >= 0; i--) x[i] = d[i] - c[i] * x[i + 1];
    }

    void bc0(int i, int j, int var, T& b0, T& c0, T& d0) {
        if (bc[id(i, j)] == BC_NOSLIP) {
            b0 = 1.0; c0 = 0.0;
            d0 = var == VAR_U ? gvx[id(i, j)] : var == VAR_V ? gvy[id(i, j)] : gT[id(i, j)];
        } else { b0 = 1.0; c0 = -1.0; d0 = 0.0; }
    }
    void bc1(int i, int j, int var, T& a1, T& b1, T& d1) {
        if (bc[id(i, j)] == BC_NOSLIP) {
            a1 = 0.0; b1 = 1.0;
            d1 = var == VAR_U ? gvx[id(i, j)] : var == VAR_V ? gvy[id(i, j)] : gT[id(i, j)];
        } else { a1 = 1.0; b1 = -1.0; d1 = 0.0; }
    }

    void build_matrix(T dt, int i, int j, int var, int dir, T* a, T* b, T* c, T* d, int n,
                      Layer2<T>& lc, Layer2<T>& tl) {
        T hx = (T)gdx, hy = (T)gdy;
        T v_vis_dx2 = v_vis / (hx * hx), t_vis_dx2 = t_vis / (hx * hx);
        T v_vis_dy2 = v_vis / (hy * hy), t_vis_dy2 = t_vis / (hy * hy);
        for (int p = 1; p < n - 1; p++) {
            if (dir == DIR_X) {
                T w = tl.U(i + p, j);
                T vis = var == VAR_T ? t_vis_dx2 : v_vis_dx2;
                a[p] = -w / (2 * hx) - vis;
                b[p] = 1 / dt + 2 * vis;
                c[p] = w / (2 * hx) - vis;
                if (var == VAR_U) d[p] = lc.U(i + p, j) / dt - v_T * tl.Tx(i + p, j);
                else if (var == VAR_V) d[p] = lc.V(i + p, j) / dt;
                else d[p] = lc.Tt(i + p, j) / dt + t_phi * tl.DissFuncX(i + p, j);
            } else {
                T w = tl.V(i, j + p);
                T vis = var == VAR_T ? t_vis_dy2 : v_vis_dy2;
                a[p] = -w / (2 * hy) - vis;
                b[p] = 1 / dt + 2 * vis;
                c[p] = w / (2 * hy) - vis;
                if (var == VAR_U) d[p] = lc.U(i, j + p) / dt;
                else if (var == VAR_V) d[p] = lc.V(i, j + p) / dt - v_T * tl.Ty(i, j + p);
                else d[p] = lc.Tt(i, j + p) / dt + t_phi * tl.DissFuncY(i, j + p);
            }
        }
    }

    void solve_direction(T dt, int num_local, std::vector<Seg2>& list, Layer2<T>& lc,
                         Layer2<T>& lt, Layer2<T>& out) {
        Layer2<T> temp_local;
        temp_local.init(nx, ny, (T)gdx, (T)gdy);
        copy_all(lt, temp_local);
        if (list.empty()) return;  // the reference reads list[0] unconditionally
        int dir = list[0].dir;
        int maxn = nx > ny ? nx : ny;
        std::vector<T> a(maxn), b(maxn), c(maxn), d(maxn), x(maxn);
        for (int it = 0; it < num_local; it++) {
            for (auto& s : list)
                for (int var = 0; var < 3; var++) {
                    int n = s.size;
                    bc0(s.px, s.py, var, b[0], c[0], d[0]);
                    build_matrix(dt, s.px, s.py, var, dir, a.data(), b.data(), c.data(), d.data(), n, lc, temp_local);
                    bc1(s.ex, s.ey, var, a[n - 1], b[n - 1], d[n - 1]);
                    thomas(a.data(), b.data(), c.data(), d.data(), x.data(), n);
                    int i = s.px, j = s.py;
                    for (int t = 0; t < n; t++) {
                        next_local.F(var, i, j) = x[t];
                        if (dir == DIR_X) i++; else j++;
                    }
                }
            if (it == 0) copy_type(next_local, temp_local, NODE_IN);
            else merge_type(next_local, temp_local, NODE_IN);
        }
        copy_type(temp_local, lt, NODE_IN);
        copy_type(next_local, out, NODE_IN);
    }

    void update_boundaries() {
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) {
                int t = GetType(i, j);
                if (t == NODE_BOUND || t == NODE_VALVE) {
                    cur.U(i, j) = gvx[id(i, j)]; cur.V(i, j) = gvy[id(i, j)]; cur.Tt(i, j) = gT[id(i, j)];
                }
            }
        copy_type(cur, next, NODE_BOUND);
        copy_type(cur, next, NODE_VALVE);
    }

    // returns 0 ok, 2 = Picard iterations exceeded MAX_GLOBAL_ITERS (100),
    // 3 = error above 10x threshold (the reference calls exit(1) for both).
    int time_step(T dt, int num_global, int num_local, double* err_out, int* iters_out) {
        create_segments();
        copy_all(cur, next);
        copy_all(cur, half);
        copy_all(cur, temp);
        int it;
        double err = div_error(next);
        int status = 0;
        for (it = 0; (it < num_global) || (err > 0.1); it++) {
            solve_direction(dt, num_local, listY, cur, temp, half);
            solve_direction(dt, num_local, listX, half, temp, next);
            err = div_error(next);
            if (it == 0) copy_type(next, temp, NODE_IN);
            else merge_type(next, temp, NODE_IN);
            if (it > 100) { status = 2; break; }
            if (err > 0.1 * 10) { status = 3; break; }
        }
        if (err_out) *err_out = err;
        if (iters_out) *iters_out = it + (status ? 1 : 0);
        if (status) return status;
        // ClearOutterCells (Solver2D.cpp:75-85)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                if (GetType(i, j) == NODE_OUT) {
                    next.U(i, j) = 0.0; next.V(i, j) = 0.0; next.Tt(i, j) = (T)startT;
                }
        copy_all(next, cur);
        return 0;
    }

    // ---- StableSolver2D -----------------------------------------------------
    Layer2<T> next_w;
    std::vector<T> q, dv;            // ScalarField2D q[0] and div
    std::vector<int> inner, bound;   // flat indices, (i, j) order
    std::vector<int> sweeps;         // relaxation sweeps of each project in the last step
    double poisson_tol = 1e-2;       // POISSON_ERR_THRESHOLD (StableSolver2D.h:23)
    int poisson_cap = 1 << 30;       // the reference has none

    void stable_init() {
        next_w.init(nx, ny, (T)gdx, (T)gdy);
        q.assign((size_t)nx * ny, (T)0);
        dv.assign((size_t)nx * ny, (T)0);
        copy_all(cur, next);
        copy_all(cur, temp);
    }

    void prepare_indices() {
        inner.clear();
        bound.clear();
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) {
                int t = GetType(i, j);
                if (t == NODE_IN) inner.push_back((int)id(i, j));
                else if (t == NODE_BOUND || t == NODE_VALVE) bound.push_back((int)id(i, j));
            }
    }

    void solve_uv(T dt) {
        Layer2<T>& tl = temp;
        for (int c : inner) {
            int i = c / ny, j = c % ny;
            next_w.U(i, j) = cur.U(i, j) + dt * (-tl.U(i, j) * tl.Ux(i, j) - tl.V(i, j) * tl.Uy(i, j) +
                                                 v_vis * (tl.Uxx(i, j) + tl.Uyy(i, j)));
        }
        for (int c : inner) {
            int i = c / ny, j = c % ny;
            next_w.V(i, j) = cur.V(i, j) + dt * (-tl.U(i, j) * tl.Vx(i, j) - tl.V(i, j) * tl.Vy(i, j) +
                                                 v_vis * (tl.Vxx(i, j) + tl.Vyy(i, j)));
        }
    }

    bool inside(int i, int j) const { return i >= 0 && j >= 0 && i < nx && j < ny; }
    bool is_in(int i, int j) { return inside(i, j) && GetType(i, j) == NODE_IN; }
    double qat(int i, int j) { return inside(i, j) ? (double)q[id(i, j)] : 0.0; }

    void relax(int c, double i0, double i1, double j0, double j1, double dx2, double dy2, double rcp,
               double& err) {
        double qn = rcp * ((i0 + i1) * dy2 + (j0 + j1) * dx2 - (double)dv[c] * dx2 * dy2);
        double ce = std::fabs((qn - (double)q[c]) / qn);
        err = (ce < err) ? err : ce;  // std::max(ce, err)
        q[c] = (T)qn;
    }

    // returns 1 when the sweep cap is reached
    int project(Layer2<T>& w, Layer2<T>& proj) {
        std::fill(dv.begin(), dv.end(), (T)0);
        for (int c : inner) dv[c] = w.Ux(c / ny, c % ny) + w.Vy(c / ny, c % ny);
        const double dx2 = gdx * gdx, dy2 = gdy * gdy, rcp = 0.5 / (dx2 + dy2);
        std::fill(q.begin(), q.end(), (T)0);
        int n = 0;
        for (;;) {
            double err = 0.0;
            for (int c : bound) {
                int i = c / ny, j = c % ny;
                double i0 = is_in(i - 1, j) ? qat(i - 1, j) : qat(i + 1, j);
                double i1 = is_in(i + 1, j) ? qat(i + 1, j) : qat(i - 1, j);
                double j0 = is_in(i, j - 1) ? qat(i, j - 1) : qat(i, j + 1);
                double j1 = is_in(i, j + 1) ? qat(i, j + 1) : qat(i, j - 1);
                relax(c, i0, i1, j0, j1, dx2, dy2, rcp, err);
            }
            for (int c : inner) {
                int i = c / ny, j = c % ny;
                relax(c, qat(i - 1, j), qat(i + 1, j), qat(i, j - 1), qat(i, j + 1), dx2, dy2, rcp, err);
            }
            n++;
            if (!(err >= poisson_tol)) break;
            if (n >= poisson_cap) { sweeps.push_back(n); return 1; }
        }
        sweeps.push_back(n);
        const T hx = (T)gdx, hy = (T)gdy;
        for (int c : inner) {
            int i = c / ny, j = c % ny;
            T gx = (q[id(i + 1, j)] - q[id(i - 1, j)]) / (2 * hx);
            T gy = (q[id(i, j + 1)] - q[id(i, j - 1)]) / (2 * hy);
            proj.U(i, j) = w.U(i, j) - gx;
            proj.V(i, j) = w.V(i, j) - gy;
        }
        return 0;
    }

    // 0 ok, 2 / 3 as time_step, 4 = relaxation sweep cap reached
    int stable_step(T dt, int num_global, double* err_out, int* iters_out) {
        copy_all(cur, temp);
        prepare_indices();
        sweeps.clear();
        double err = div_error(next);
        int it, status = 0;
        for (it = 0; (it < num_global) || (err > 0.1); it++) {
            copy_all(cur, next_w);
            solve_uv(dt);
            if (project(next_w, next)) { status = 4; break; }
            err = div_error(next);
            merge_type(next, temp, NODE_IN);
            if (it > 100) { status = 2; break; }
            if (err > 0.1 * 10) { status = 3; break; }
        }
        if (err_out) *err_out = err;
        if (iters_out) *iters_out = it + (status ? 1 : 0);
        if (status) return status;
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                if (GetType(i, j) == NODE_OUT) {
                    next.U(i, j) = 0.0; next.V(i, j) = 0.0; next.Tt(i, j) = (T)startT;
                }
        copy_all(next, cur);
        return 0;
    }
};

struct H2 { int dtype; void* s; int method = 1; };

}  // namespace orc2

using namespace orc2;

#define DISPATCH2(h, ...)                                                              \
    do {                                                                                \
        if (((H2*)(h))->dtype == 0) { auto* s = (Solver2<float>*)((H2*)(h))->s; __VA_ARGS__; } \
        else { auto* s = (Solver2<double>*)((H2*)(h))->s; __VA_ARGS__; }                \
    } while (0)

extern "C" {

void* orc2_create(int dtype, int nx, int ny, double dx, double dy, double startT,
                  const uint8_t* type, const uint8_t* bc, const double* vx, const double* vy,
                  const double* T, double v_T, double v_vis, double t_vis, double t_phi) {
    H2* h = new H2;
    h->dtype = dtype;
    if (dtype == 0) {
        auto* s = new Solver2<float>;
        s->init(nx, ny, dx, dy, startT, type, bc, vx, vy, T);
        s->v_T = (float)v_T; s->v_vis = (float)v_vis; s->t_vis = (float)t_vis; s->t_phi = (float)t_phi;
        h->s = s;
    } else {
        auto* s = new Solver2<double>;
        s->init(nx, ny, dx, dy, startT, type, bc, vx, vy, T);
        s->v_T = v_T; s->v_vis = v_vis; s->t_vis = t_vis; s->t_phi = t_phi;
        h->s = s;
    }
    return h;
}

void orc2_destroy(void* h) {
    DISPATCH2(h, delete s);
    delete (H2*)h;
}

void orc2_update_boundaries(void* h) { DISPATCH2(h, s->update_boundaries()); }

int orc2_step(void* h, double dt, int G, int L, double* err, int* iters) {
    int r = 0;
    if (((H2*)h)->method == 2) DISPATCH2(h, r = s->stable_step((decltype(s->v_T))dt, G, err, iters));
    else DISPATCH2(h, r = s->time_step((decltype(s->v_T))dt, G, L, err, iters));
    return r;
}

// Grid2D::Prepare(t) between steps (FluidSolver2D.cpp:131): new cell types, BC
// kinds and node values (UpdateBoundaries and the next step read them; both
// solvers rebuild their lists from the types every step).
void orc2_set_grid(void* h, const uint8_t* type, const uint8_t* bc, const double* vx, const double* vy,
                   const double* T) {
    DISPATCH2(h, {
        using TT = decltype(s->v_T);
        const size_t N = s->type.size();
        s->type.assign(type, type + N);
        s->bc.assign(bc, bc + N);
        for (size_t i = 0; i < N; i++) {
            s->gvx[i] = (TT)vx[i];
            s->gvy[i] = (TT)vy[i];
            s->gT[i] = (TT)T[i];
        }
    });
}

// SolveU / SolveV alone (test hook): out = cur + dt * (advection-diffusion of tl)
// on NODE_IN cells; other cells of out as given.
void orc2_stable_solve_uv(void* h, double dt, const double* cu, const double* cv, const double* tu,
                          const double* tv, double* ou, double* ov) {
    DISPATCH2(h, {
        using TT = decltype(s->v_T);
        const size_t N = s->type.size();
        for (size_t i = 0; i < N; i++) {
            s->cur.u[i] = (TT)cu[i]; s->cur.v[i] = (TT)cv[i];
            s->temp.u[i] = (TT)tu[i]; s->temp.v[i] = (TT)tv[i];
            s->next_w.u[i] = (TT)ou[i]; s->next_w.v[i] = (TT)ov[i];
        }
        s->prepare_indices();
        s->solve_uv((TT)dt);
        for (size_t i = 0; i < N; i++) { ou[i] = s->next_w.u[i]; ov[i] = s->next_w.v[i]; }
    });
}

// Config.h solver ids: 1 = ADI, 2 = Stable (projection)
void orc2_set_method(void* h, int method) {
    ((H2*)h)->method = method;
    if (method == 2) DISPATCH2(h, s->stable_init());
}

// relaxation settings of the Stable solver (the reference's: 1e-2, no cap)
void orc2_set_poisson(void* h, double tol, int cap) {
    DISPATCH2(h, { s->poisson_tol = tol; s->poisson_cap = cap; });
}

// relaxation sweeps of each project of the last step; returns their count
int orc2_poisson_sweeps(void* h, int32_t* out, int max) {
    int n = 0;
    DISPATCH2(h, {
        n = (int)s->sweeps.size();
        for (int k = 0; k < n && k < max; k++) out[k] = s->sweeps[k];
    });
    return n;
}

// one project of the Stable solver on a given intermediate field w (all cells):
// q, the projected u and v (NODE_IN cells; others as the layer held them) and
// the sweep count.  Returns 1 at the sweep cap.
int orc2_stable_project(void* h, const double* wu, const double* wv, double* q_out, double* pu,
                        double* pv, int* sweeps) {
    int r = 0;
    DISPATCH2(h, {
        using TT = decltype(s->v_T);
        s->prepare_indices();
        s->sweeps.clear();
        for (size_t i = 0; i < s->next_w.u.size(); i++) {
            s->next_w.u[i] = (TT)wu[i];
            s->next_w.v[i] = (TT)wv[i];
        }
        r = s->project(s->next_w, s->next);
        for (size_t i = 0; i < s->q.size(); i++) {
            q_out[i] = s->q[i];
            pu[i] = s->next.u[i];
            pv[i] = s->next.v[i];
        }
        if (sweeps) *sweeps = s->sweeps.empty() ? 0 : s->sweeps.back();
    });
    return r;
}

// which: 0 = cur, 1 = next
void orc2_get_layer_raw(void* h, int which, double* u, double* v, double* T) {
    DISPATCH2(h, {
        auto& L = which == 0 ? s->cur : s->next;
        for (size_t i = 0; i < L.u.size(); i++) { u[i] = L.u[i]; v[i] = L.v[i]; T[i] = L.t[i]; }
    });
}

// axis 0: listY (X-direction segments, one per row j); axis 1: listX (Y-direction, one per column i)
int orc2_segments(void* h, int axis, int32_t* out) {
    int n = 0;
    DISPATCH2(h, {
        s->create_segments();
        auto& L = axis == 0 ? s->listY : s->listX;
        n = (int)L.size();
        if (out)
            for (int q = 0; q < n; q++) {
                out[5 * q + 0] = L[q].px; out[5 * q + 1] = L[q].py; out[5 * q + 2] = L[q].ex;
                out[5 * q + 3] = L[q].ey; out[5 * q + 4] = L[q].size;
            }
    });
    return n;
}

}  // extern "C"
