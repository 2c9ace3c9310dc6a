// ORACLE -- test infrastructure only.
//
// CPU restatement of the reference 3D ADI time step (samech-git/cmc-fluid-solver,
// CPU backend).  It is the checker the HIP product is compared against; only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
// Nothing in the product links it.
//
// Templated on the field type T (float or double), mirroring the reference's
// FTYPE switch (src/Common/Geometry.h:21).
//
// Follows, function by function:
//   gen_segments          src/FluidSolver3D/Grid3D.cpp:47-127   (GenerateListSegments)
//   solve_tridiagonal     src/Common/Algorithms.h:21-38         (SolveTridiagonal)
//   apply_bc0 / apply_bc1 src/FluidSolver3D/AdiSolver3D.cpp:804-852
//   build_matrix          src/FluidSolver3D/AdiSolver3D.cpp:732-802
//   d_x/d_y/d_z, diss_*   src/FluidSolver3D/TimeLayer3D.h:338-343, 554-588
//   merge                 src/FluidSolver3D/TimeLayer3D.h:415-436, 664-683
//   solve_direction       src/FluidSolver3D/AdiSolver3D.cpp:564-666 (CPU branch)
//   time_step             src/FluidSolver3D/AdiSolver3D.cpp:306-391
//   update_boundaries     src/FluidSolver3D/AdiSolver3D.cpp:286-304 (CPU branch),
//                         src/FluidSolver3D/TimeLayer3D.h:926-951
//   div_error             src/FluidSolver3D/TimeLayer3D.h:595-641
//   get_layer             src/FluidSolver3D/Solver3D.cpp:21-25, TimeLayer3D.h:819-924
//   set_boundary          src/FluidSolver3D/Grid3D.h:121 (Grid3D::SetData)
//
// Merge modes:
//   0 = CPU semantics (AdiSolver3D.cpp:591-604, 634-655, 347-357): the 4 fields
//       of a sweep are solved against one temp layer, then temp <- (temp+out)/2
//       on NODE_IN; one extra merge after each X sweep of a global iteration.
//   1 = reference-GPU ordering (AdiSolver3D.h:21 INTERNAL_MERGE_ENABLE=1,
//       AdiSolver3D.cu:337-364, 479-486): U, V, W, T solved one after another,
//       each merging its own field into temp right after its solve; no host merges.
//
// Layers other than the state: the reference allocates next, half and temp
// without initialising them (TimeLayer3D.h:1065-1075), so their values at nodes
// that no segment covers are unspecified.  Here every layer starts as a copy of
// the state, and set_fields writes every layer: such nodes keep their values.
//
// Shared segment endpoints (a one-cell wall between two runs on one line) are
// written by both segments in the reference, in parallel (a race when the wall
// is BC_FREE, AdiSolver3D.cpp:593-603).  Here the solve runs in parallel into
// private buffers and the endpoints are scattered serially in list order, so
// the later segment of the list wins, independent of the thread count.
//
// What pins each function (tests/test_oracle_pins.py, test_oracle_adjoint.py):
//   gen_segments        numpy run-length brute force, known answers, invariants,
//                       Z-block permutation
//   solve_tridiagonal   bit-exact against the reference's Algorithms.h compiled
//                       from the reference tree (oracle/_ref), and a dense solve
//   FluidParams         bit-exact against the reference's Geometry.h (oracle/_ref)
//   build_matrix, BCs,  closed forms of single steps (diffusion eigenmode decay,
//   stencils, diss_*    pressure and advection signs per axis, every dissipation
//                       term) and a dense brute force of whole steps
//                       (tests/brute_adi.py, written from the model)
//   time_step, merge,   dense brute force of whole steps, both merge modes;
//   update_boundaries   zero state stays zero
//   div_error           closed form (linear field)
//   get_layer           closed form (index-valued state, NODE_OUT fill)
//   Adjoint             central differences of time_step in every input group

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

enum { NODE_IN = 0, NODE_OUT = 1, NODE_BOUND = 2, NODE_VALVE = 3 };
enum { BC_NOSLIP = 0, BC_FREE = 1 };
enum { AX_X = 0, AX_Y = 1, AX_Z = 2 };
enum { VAR_U = 0, VAR_V = 1, VAR_W = 2, VAR_T = 3 };

struct Seg {
    int32_t px, py, pz, ex, ey, ez, size;
};

// Grid3D::GenerateListSegments (Grid3D.cpp:47-127).  dim2/dim3 loop order and
// the Z-block reordering (nblockZ > 1, X and Y only) follow the reference.
static void gen_segments(const uint8_t* type, int nx, int ny, int nz, int axis,
                         int nblockZ, std::vector<Seg>& out) {
    out.clear();
    int dim2, dim3;
    switch (axis) {
        case AX_X: dim2 = ny; dim3 = nz; break;
        case AX_Y: dim2 = nx; dim3 = nz; break;
        default:   dim2 = nx; dim3 = ny; break;
    }
    int nb = (axis == AX_Z) ? 1 : nblockZ;
    if (nb < 1) nb = 1;
    int block_start = 0;
    int block_end = dim3 / nb;
    auto type_at = [&](int i, int j, int k) {
        return type[((size_t)i * ny + j) * nz + k];
    };
    while (block_end <= dim3) {
        for (int i = 0; i < dim2; i++)
            for (int j = block_start; j < block_end; j++) {
                int px, py, pz, ix = 0, iy = 0, iz = 0;
                switch (axis) {
                    case AX_X: px = 0; py = i; pz = j; ix = 1; break;
                    case AX_Y: px = i; py = 0; pz = j; iy = 1; break;
                    default:   px = i; py = j; pz = 0; iz = 1; break;
                }
                int state = 0;
                Seg ns{};
                while (px + ix < nx && py + iy < ny && pz + iz < nz) {
                    if (type_at(px + ix, py + iy, pz + iz) == NODE_IN) {
                        if (state == 0) { ns.px = px; ns.py = py; ns.pz = pz; }
                        state = 1;
                    } else if (state == 1) {
                        ns.ex = px + ix; ns.ey = py + iy; ns.ez = pz + iz;
                        ns.size = (ns.ex - ns.px) + (ns.ey - ns.py) + (ns.ez - ns.pz) + 1;
                        out.push_back(ns);
                        state = 0;
                    }
                    px += ix; py += iy; pz += iz;
                }
            }
        block_start = block_end;
        block_end += (block_end + dim3 / nb == (dim3 / nb) * nb) ? dim3 / nb + dim3 % nb : dim3 / nb;
    }
}

// Algorithms.h:21-38, same operation order.
template <class T>
static void solve_tridiagonal(T* a, T* b, T* c, T* d, T* x, int num) {
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

template <class T>
struct Layer {
    int nx = 0, ny = 0, nz = 0;
    size_t N = 0, halo = 0;
    T dx = 0, dy = 0, dz = 0;
    std::vector<T> f[4];
    void init(int x, int y, int z, T hx, T hy, T hz) {
        nx = x; ny = y; nz = z; dx = hx; dy = hy; dz = hz;
        N = (size_t)x * y * z;
        halo = (size_t)y * z;  // one x-plane each side, like the reference's cur layer
        for (auto& v : f) v.assign(N + 2 * halo, (T)0);
    }
    inline T& e(int var, int i, int j, int k) {
        return f[var][halo + ((size_t)i * ny + j) * nz + k];
    }
    // ScalarField3D::d_x/d_y/d_z (TimeLayer3D.h:338-340)
    inline T d_x(int var, int i, int j, int k) { return (e(var, i + 1, j, k) - e(var, i - 1, j, k)) / (2 * dx); }
    inline T d_y(int var, int i, int j, int k) { return (e(var, i, j + 1, k) - e(var, i, j - 1, k)) / (2 * dy); }
    inline T d_z(int var, int i, int j, int k) { return (e(var, i, j, k + 1) - e(var, i, j, k - 1)) / (2 * dz); }
    // TimeLayer3D::DissFuncX/Y/Z (TimeLayer3D.h:554-588)
    inline T diss_x(int i, int j, int k) {
        T u_x = d_x(VAR_U, i, j, k), v_x = d_x(VAR_V, i, j, k), w_x = d_x(VAR_W, i, j, k);
        T u_y = d_y(VAR_U, i, j, k), u_z = d_z(VAR_U, i, j, k);
        return 2 * u_x * u_x + v_x * v_x + w_x * w_x + v_x * u_y + w_x * u_z;
    }
    inline T diss_y(int i, int j, int k) {
        T u_y = d_y(VAR_U, i, j, k), v_y = d_y(VAR_V, i, j, k), w_y = d_y(VAR_W, i, j, k);
        T v_x = d_x(VAR_V, i, j, k), v_z = d_z(VAR_V, i, j, k);
        return u_y * u_y + 2 * v_y * v_y + w_y * w_y + u_y * v_x + w_y * v_z;
    }
    inline T diss_z(int i, int j, int k) {
        T u_z = d_z(VAR_U, i, j, k), v_z = d_z(VAR_V, i, j, k), w_z = d_z(VAR_W, i, j, k);
        T w_x = d_x(VAR_W, i, j, k), w_y = d_y(VAR_W, i, j, k);
        return u_z * u_z + v_z * v_z + 2 * w_z * w_z + u_z * w_x + v_z * w_y;
    }
};

template <class T>
struct Solver {
    int nx, ny, nz;
    size_t N;
    T dx, dy, dz;
    // grid (Grid3D Node array as SoA; Node.v and Node.T are FTYPE)
    std::vector<uint8_t> type, bc_vel, bc_temp;
    std::vector<T> gv[3], gT;
    // FluidParams (Geometry.h:538-562): stored as FTYPE
    T v_T, v_vis, t_vis, t_phi;
    int merge_mode = 0;
    int nblockZ = 1;
    std::vector<Seg> segs[3];
    Layer<T> L_cur, L_next, L_half, L_temp;
    Layer<T>*cur, *next, *half, *temp;
    double diffError = 0.0;
    int max_n;

    void init(int x, int y, int z, double hx, double hy, double hz, const uint8_t* t,
              const uint8_t* bv, const uint8_t* bt, const double* vx, const double* vy,
              const double* vz, const double* tT) {
        nx = x; ny = y; nz = z;
        N = (size_t)x * y * z;
        dx = (T)hx; dy = (T)hy; dz = (T)hz;
        type.assign(t, t + N);
        bc_vel.assign(bv, bv + N);
        bc_temp.assign(bt, bt + N);
        for (int c = 0; c < 3; c++) gv[c].resize(N);
        gT.resize(N);
        for (size_t i = 0; i < N; i++) {
            gv[0][i] = (T)vx[i]; gv[1][i] = (T)vy[i]; gv[2][i] = (T)vz[i]; gT[i] = (T)tT[i];
        }
        L_cur.init(x, y, z, dx, dy, dz);
        L_next.init(x, y, z, dx, dy, dz);
        L_half.init(x, y, z, dx, dy, dz);
        L_temp.init(x, y, z, dx, dy, dz);
        cur = &L_cur; next = &L_next; half = &L_half; temp = &L_temp;
        // TimeLayer3D(backend, grid) -> CopyFromGrid(grid) (TimeLayer3D.h:734-781)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                for (int k = 0; k < nz; k++) {
                    size_t id = idx(i, j, k);
                    cur->e(VAR_U, i, j, k) = gv[0][id];
                    cur->e(VAR_V, i, j, k) = gv[1][id];
                    cur->e(VAR_W, i, j, k) = gv[2][id];
                    cur->e(VAR_T, i, j, k) = gT[id];
                }
        copy_state_to_all();
        max_n = std::max(nx, std::max(ny, nz));
    }
    void copy_state_to_all() {
        for (Layer<T>* L : {next, half, temp})
            for (int v = 0; v < 4; v++) L->f[v] = cur->f[v];
    }
    inline size_t idx(int i, int j, int k) const { return ((size_t)i * ny + j) * nz + k; }

    void create_segments() {
        // AdiSolver3D::CreateSegments (AdiSolver3D.cpp:553-562): X, Y, Z.
        gen_segments(type.data(), nx, ny, nz, AX_X, nblockZ, segs[AX_X]);
        gen_segments(type.data(), nx, ny, nz, AX_Y, nblockZ, segs[AX_Y]);
        gen_segments(type.data(), nx, ny, nz, AX_Z, nblockZ, segs[AX_Z]);
    }

    // ApplyBC0 (AdiSolver3D.cpp:804-827)
    void apply_bc0(int i, int j, int k, int var, T& b0, T& c0, T& d0) {
        size_t id = idx(i, j, k);
        if ((var == VAR_T && bc_temp[id] == BC_FREE) || (var != VAR_T && bc_vel[id] == BC_FREE)) {
            b0 = 2.0; c0 = -1.0; d0 = 0.0;
        } else {
            b0 = 1.0; c0 = 0.0;
            d0 = (var == VAR_T) ? gT[id] : gv[var][id];
        }
    }
    // ApplyBC1 (AdiSolver3D.cpp:829-852)
    void apply_bc1(int i, int j, int k, int var, T& a1, T& b1, T& d1) {
        size_t id = idx(i, j, k);
        if ((var == VAR_T && bc_temp[id] == BC_FREE) || (var != VAR_T && bc_vel[id] == BC_FREE)) {
            a1 = -1.0; b1 = 2.0; d1 = 0.0;
        } else {
            a1 = 0.0; b1 = 1.0;
            d1 = (var == VAR_T) ? gT[id] : gv[var][id];
        }
    }

    // BuildMatrix (AdiSolver3D.cpp:732-802), rows p = 1..n-2.
    void build_matrix(T dt, int i, int j, int k, int var, int dir, T* a, T* b, T* c, T* d,
                      int n, Layer<T>* lc, Layer<T>* lt) {
        T vis_dx2, vis_dy2, vis_dz2;
        T hx = lc->dx, hy = lc->dy, hz = lc->dz;
        if (var != VAR_T) {
            vis_dx2 = v_vis / (hx * hx); vis_dy2 = v_vis / (hy * hy); vis_dz2 = v_vis / (hz * hz);
        } else {
            vis_dx2 = t_vis / (hx * hx); vis_dy2 = t_vis / (hy * hy); vis_dz2 = t_vis / (hz * hz);
        }
        for (int p = 1; p < n - 1; p++) {
            switch (dir) {
                case AX_X: {
                    T w = lt->e(VAR_U, i + p, j, k);
                    a[p] = -w / (2 * hx) - vis_dx2;
                    b[p] = 3 / dt + 2 * vis_dx2;
                    c[p] = w / (2 * hx) - vis_dx2;
                    switch (var) {
                        case VAR_U: d[p] = lc->e(VAR_U, i + p, j, k) * 3 / dt - v_T * lt->d_x(VAR_T, i + p, j, k); break;
                        case VAR_V: d[p] = lc->e(VAR_V, i + p, j, k) * 3 / dt; break;
                        case VAR_W: d[p] = lc->e(VAR_W, i + p, j, k) * 3 / dt; break;
                        default:    d[p] = lc->e(VAR_T, i + p, j, k) * 3 / dt + t_phi * lt->diss_x(i + p, j, k); break;
                    }
                    break;
                }
                case AX_Y: {
                    T w = lt->e(VAR_V, i, j + p, k);
                    a[p] = -w / (2 * hy) - vis_dy2;
                    b[p] = 3 / dt + 2 * vis_dy2;
                    c[p] = w / (2 * hy) - vis_dy2;
                    switch (var) {
                        case VAR_U: d[p] = lc->e(VAR_U, i, j + p, k) * 3 / dt; break;
                        case VAR_V: d[p] = lc->e(VAR_V, i, j + p, k) * 3 / dt - v_T * lt->d_y(VAR_T, i, j + p, k); break;
                        case VAR_W: d[p] = lc->e(VAR_W, i, j + p, k) * 3 / dt; break;
                        default:    d[p] = lc->e(VAR_T, i, j + p, k) * 3 / dt + t_phi * lt->diss_y(i, j + p, k); break;
                    }
                    break;
                }
                default: {
                    T w = lt->e(VAR_W, i, j, k + p);
                    a[p] = -w / (2 * hz) - vis_dz2;
                    b[p] = 3 / dt + 2 * vis_dz2;
                    c[p] = w / (2 * hz) - vis_dz2;
                    switch (var) {
                        case VAR_U: d[p] = lc->e(VAR_U, i, j, k + p) * 3 / dt; break;
                        case VAR_V: d[p] = lc->e(VAR_V, i, j, k + p) * 3 / dt; break;
                        case VAR_W: d[p] = lc->e(VAR_W, i, j, k + p) * 3 / dt - v_T * lt->d_z(VAR_T, i, j, k + p); break;
                        default:    d[p] = lc->e(VAR_T, i, j, k + p) * 3 / dt + t_phi * lt->diss_z(i, j, k + p); break;
                    }
                    break;
                }
            }
        }
    }

    // SolveSegment (AdiSolver3D.cpp:687-705) without the scatter: x[0..n-1].
    void solve_segment(T dt, const Seg& s, int var, int dir, Layer<T>* lc, Layer<T>* lt, T* a,
                       T* b, T* c, T* d, T* x) {
        int n = s.size;
        apply_bc0(s.px, s.py, s.pz, var, b[0], c[0], d[0]);
        apply_bc1(s.ex, s.ey, s.ez, var, a[n - 1], b[n - 1], d[n - 1]);
        build_matrix(dt, s.px, s.py, s.pz, var, dir, a, b, c, d, n, lc, lt);
        solve_tridiagonal(a, b, c, d, x, n);
    }

    inline void seg_cell(const Seg& s, int dir, int t, int& i, int& j, int& k) {
        i = s.px; j = s.py; k = s.pz;
        if (dir == AX_X) i += t; else if (dir == AX_Y) j += t; else k += t;
    }

    // MergeLayerTo(grid, temp, NODE_IN) (TimeLayer3D.h:664-683, 415-436)
    void merge(Layer<T>* src, Layer<T>* dst) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                for (int k = 0; k < nz; k++)
                    if (type[idx(i, j, k)] == NODE_IN)
                        for (int v = 0; v < 4; v++)
                            dst->e(v, i, j, k) = (dst->e(v, i, j, k) + src->e(v, i, j, k)) / 2;
    }

    // Solve all segments of one sweep for the variables in [v0, v1) into out,
    // reading lc (RHS) and lt (non-linear layer).  merge_var >= 0 merges that
    // variable into lt right after the scatter (reference-GPU ordering).
    void sweep(int dir, T dt, Layer<T>* lc, Layer<T>* lt, Layer<T>* out, int v0, int v1) {
        std::vector<Seg>& list = segs[dir];
        int ns = (int)list.size();
        int nv = v1 - v0;
        // per (segment, var) private solution buffers
        std::vector<size_t> off(ns + 1, 0);
        for (int s = 0; s < ns; s++) off[s + 1] = off[s] + (size_t)list[s].size;
        std::vector<T> X(off[ns] * nv);
#pragma omp parallel
        {
            std::vector<T> a(max_n + 2), b(max_n + 2), c(max_n + 2), d(max_n + 2);
#pragma omp for schedule(static)
            for (int s = 0; s < ns; s++)
                for (int v = v0; v < v1; v++)
                    solve_segment(dt, list[s], v, dir, lc, lt, a.data(), b.data(), c.data(),
                                  d.data(), &X[off[s] * nv + (size_t)(v - v0) * list[s].size]);
        }
        // scatter: interior cells in parallel (each NODE_IN cell belongs to one
        // segment per axis), endpoints serially in list order.
#pragma omp parallel for schedule(static)
        for (int s = 0; s < ns; s++) {
            const Seg& sg = list[s];
            for (int t = 1; t < sg.size - 1; t++) {
                int i, j, k;
                seg_cell(sg, dir, t, i, j, k);
                for (int v = v0; v < v1; v++)
                    out->e(v, i, j, k) = X[off[s] * nv + (size_t)(v - v0) * sg.size + t];
            }
        }
        for (int s = 0; s < ns; s++) {
            const Seg& sg = list[s];
            int ts[2] = {0, sg.size - 1};
            for (int q = 0; q < 2; q++) {
                int i, j, k;
                seg_cell(sg, dir, ts[q], i, j, k);
                for (int v = v0; v < v1; v++)
                    out->e(v, i, j, k) = X[off[s] * nv + (size_t)(v - v0) * sg.size + ts[q]];
            }
        }
    }

    // Reference-GPU in-kernel merge of one variable on a sweep's segment cells
    // (AdiSolver3D.cu:337-364): temp.var <- (temp.var + x)/2 on NODE_IN cells.
    void merge_var_on_segments(int dir, int var, Layer<T>* out, Layer<T>* lt) {
        std::vector<Seg>& list = segs[dir];
        int ns = (int)list.size();
#pragma omp parallel for schedule(static)
        for (int s = 0; s < ns; s++) {
            const Seg& sg = list[s];
            for (int t = 0; t < sg.size; t++) {
                int i, j, k;
                seg_cell(sg, dir, t, i, j, k);
                if (type[idx(i, j, k)] == NODE_IN)
                    lt->e(var, i, j, k) = (lt->e(var, i, j, k) + out->e(var, i, j, k)) / 2;
            }
        }
    }

    // SolveDirection (AdiSolver3D.cpp:564-666)
    void solve_direction(int dir, T dt, int num_local, Layer<T>* lc, Layer<T>* lt, Layer<T>* out) {
        for (int it = 0; it < num_local; it++) {
            if (merge_mode == 0) {
                sweep(dir, dt, lc, lt, out, 0, 4);
                merge(out, lt);
            } else {
                for (int v = 0; v < 4; v++) {
                    sweep(dir, dt, lc, lt, out, v, v + 1);
                    merge_var_on_segments(dir, v, out, lt);
                }
            }
        }
    }

    // CopyLayerTo(grid, dest, type) (TimeLayer3D.h:726-732, 394-413)
    void copy_typed(Layer<T>* src, Layer<T>* dst, int t) {
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                for (int k = 0; k < nz; k++)
                    if (type[idx(i, j, k)] == t)
                        for (int v = 0; v < 4; v++) dst->e(v, i, j, k) = src->e(v, i, j, k);
    }

    // UpdateBoundaries, CPU branch (AdiSolver3D.cpp:290-294; TimeLayer3D.h:926-951)
    void update_boundaries() {
        for (int t : {NODE_BOUND, NODE_VALVE})
            for (int i = 0; i < nx; i++)
                for (int j = 0; j < ny; j++)
                    for (int k = 0; k < nz; k++) {
                        size_t id = idx(i, j, k);
                        if (type[id] == t) {
                            cur->e(VAR_U, i, j, k) = gv[0][id];
                            cur->e(VAR_V, i, j, k) = gv[1][id];
                            cur->e(VAR_W, i, j, k) = gv[2][id];
                            cur->e(VAR_T, i, j, k) = gT[id];
                        }
                    }
    }

    // EvalDivError (TimeLayer3D.h:595-641), single rank: i < dimx-1.
    double div_error(Layer<T>* L) {
        double err = 0.0;
        int count = 0;
        T hx = L->dx, hy = L->dy, hz = L->dz;
        for (int i = 0; i < nx - 1; i++)
            for (int j = 0; j < ny - 1; j++)
                for (int k = 0; k < nz - 1; k++)
                    if (type[idx(i, j, k)] == NODE_IN) {
                        double err_x = (L->e(VAR_U, i, j, k) + L->e(VAR_U, i, j - 1, k) + L->e(VAR_U, i, j - 1, k - 1) + L->e(VAR_U, i, j, k - 1) -
                                        L->e(VAR_U, i - 1, j, k) - L->e(VAR_U, i - 1, j - 1, k) - L->e(VAR_U, i - 1, j - 1, k - 1) - L->e(VAR_U, i - 1, j, k - 1)) * hz * hy / 4.0;
                        double err_y = (L->e(VAR_V, i, j, k) + L->e(VAR_V, i - 1, j, k) + L->e(VAR_V, i - 1, j, k - 1) + L->e(VAR_V, i, j, k - 1) -
                                        L->e(VAR_V, i, j - 1, k) - L->e(VAR_V, i - 1, j - 1, k) - L->e(VAR_V, i - 1, j - 1, k - 1) - L->e(VAR_V, i, j - 1, k - 1)) * hx * hz / 4.0;
                        double err_z = (L->e(VAR_W, i, j, k) + L->e(VAR_W, i, j - 1, k) + L->e(VAR_W, i - 1, j - 1, k) + L->e(VAR_W, i - 1, j, k) -
                                        L->e(VAR_W, i, j, k - 1) - L->e(VAR_W, i, j - 1, k - 1) - L->e(VAR_W, i - 1, j - 1, k - 1) - L->e(VAR_W, i - 1, j, k - 1)) * hx * hy / 4.0;
                        err += std::fabs(err_x + err_y + err_z);
                        count++;
                    }
        return err / count;
    }

    // TimeStep (AdiSolver3D.cpp:306-391).  Returns 1 when the divergence guard
    // fires (the reference throws before the swap, :371-374, 388-390).
    int time_step(T dt, int num_global, int num_local, bool compute_error, double* err_out) {
        copy_typed(cur, next, NODE_BOUND);
        copy_typed(cur, next, NODE_VALVE);
        for (int v = 0; v < 4; v++) temp->f[v] = cur->f[v];  // CopyLayerTo(temp)
        for (int it = 0; it < num_global; it++) {
            solve_direction(AX_Z, dt, num_local, cur, temp, next);
            solve_direction(AX_Y, dt, num_local, next, temp, half);
            solve_direction(AX_X, dt, num_local, half, temp, next);
            if (merge_mode == 0) merge(next, temp);
        }
        if (compute_error) diffError = div_error(next);
        if (err_out) *err_out = diffError;
        if (diffError > 0.01) return 1;
        std::swap(cur, next);
        return 0;
    }
};

// ---------------------------------------------------------------------------
// Discrete adjoint of TimeStep (CPU merge order, num_local = 1): the transpose
// of every linear step of the forward map, applied in reverse.  Per sweep and
// segment the forward solves M(temp) x = d(in, temp); its adjoint solves
// M^T lam = xbar (+ the merge's share of tempbar) and scatters
//   inbar = (3/dt) lam,  Mbar = -lam x^T  -> tempbar (advecting velocity) and
//   the viscosity gradients, dbar -> tempbar through the pressure and
//   dissipation stencils, and lam on Dirichlet end rows -> node-value gradients.
// The Picard layer is merged as temp <- (temp + x)/2 (twice after X), so its
// adjoint scales tempbar by 1/2 (1/4) and feeds 1/2 (3/4) into xbar.
// Node values enter through UpdateBoundaries (state and Picard layer on
// BOUND/VALVE nodes) and the Dirichlet end rows.  Shared BC_FREE end nodes
// are not supported (the forward's write order there is arbitrary).
template <class T>
struct Adjoint {
    Solver<T>& s;
    size_t N;
    int G = 0;
    T dt = 0;
    struct Tape {
        std::vector<T> S;         // state after UpdateBoundaries, [4][N]
        std::vector<T> tin, out;  // per sweep k = 3*it + {0: Z, 1: Y, 2: X}: [3G][4][N]
    };
    std::vector<Tape> tape;
    Layer<T> lt;                  // Picard layer of the sweep being reversed (with halo)

    explicit Adjoint(Solver<T>& sv) : s(sv), N(sv.N) { lt.init(s.nx, s.ny, s.nz, s.dx, s.dy, s.dz); }

    void save(Layer<T>* L, std::vector<T>& dst, size_t slot) {
        for (int f = 0; f < 4; f++)
            std::memcpy(&dst[(slot * 4 + f) * N], &L->f[f][L->halo], sizeof(T) * N);
    }

    // TimeStep (merge mode 0, L = 1) keeping the layers the adjoint needs.
    void record_step(T dt_, int G_) {
        dt = dt_;
        G = G_;
        Tape tp;
        tp.S.resize(4 * N);
        tp.tin.resize((size_t)3 * G * 4 * N);
        tp.out.resize((size_t)3 * G * 4 * N);
        save(s.cur, tp.S, 0);
        s.copy_typed(s.cur, s.next, NODE_BOUND);
        s.copy_typed(s.cur, s.next, NODE_VALVE);
        for (int v = 0; v < 4; v++) s.temp->f[v] = s.cur->f[v];
        for (int it = 0; it < G; it++) {
            Layer<T>* ins[3] = {s.cur, s.next, s.half};
            Layer<T>* outs[3] = {s.next, s.half, s.next};
            const int dirs[3] = {AX_Z, AX_Y, AX_X};
            for (int q = 0; q < 3; q++) {
                size_t k = (size_t)3 * it + q;
                save(s.temp, tp.tin, k);
                s.sweep(dirs[q], dt, ins[q], s.temp, outs[q], 0, 4);
                s.merge(outs[q], s.temp);
                save(outs[q], tp.out, k);
            }
            s.merge(s.next, s.temp);
        }
        std::swap(s.cur, s.next);
        tape.push_back(std::move(tp));
    }

    inline bool in_node(size_t i) const { return s.type[i] == NODE_IN; }

    // Reverses one sweep.  xb: adjoint of the sweep output (NODE_IN nodes);
    // tb: adjoint of the Picard layer after the sweep, replaced by the adjoint
    // before it; inb receives the adjoint of the sweep input.
    void adj_sweep(int dir, const T* tinL, const T* outL, const std::vector<T>& xb,
                   std::vector<T>& tb, T alpha, T beta, std::vector<T>& inb, std::vector<T>& gb,
                   double* pb) {
        for (int f = 0; f < 4; f++) std::memcpy(&lt.f[f][lt.halo], tinL + f * N, sizeof(T) * N);
        std::vector<Seg>& list = s.segs[dir];
        std::vector<T> lam(4 * N, (T)0);
        const int mx = s.max_n + 2;
        std::vector<T> a(mx), b(mx), c(mx), d(mx), A(mx), B(mx), C(mx), R(mx), X(mx);
        for (const Seg& sg : list) {
            const int n = sg.size;
            for (int var = 0; var < 4; var++) {
                std::fill(a.begin(), a.end(), (T)0);
                std::fill(c.begin(), c.end(), (T)0);
                s.apply_bc0(sg.px, sg.py, sg.pz, var, b[0], c[0], d[0]);
                s.apply_bc1(sg.ex, sg.ey, sg.ez, var, a[n - 1], b[n - 1], d[n - 1]);
                s.build_matrix(dt, sg.px, sg.py, sg.pz, var, dir, a.data(), b.data(), c.data(),
                               d.data(), n, &lt, &lt);
                for (int t = 0; t < n; t++) {
                    A[t] = t > 0 ? c[t - 1] : (T)0;
                    B[t] = b[t];
                    C[t] = t < n - 1 ? a[t + 1] : (T)0;
                    int i, j, k;
                    s.seg_cell(sg, dir, t, i, j, k);
                    size_t id = s.idx(i, j, k);
                    R[t] = in_node(id) ? xb[var * N + id] + beta * tb[var * N + id] : (T)0;
                }
                solve_tridiagonal(A.data(), B.data(), C.data(), R.data(), X.data(), n);
                for (int t = 0; t < n; t++) {
                    int i, j, k;
                    s.seg_cell(sg, dir, t, i, j, k);
                    size_t id = s.idx(i, j, k);
                    if (t > 0 && t < n - 1) {
                        lam[var * N + id] = X[t];
                    } else {
                        // end rows: x = node value unless BC_FREE
                        bool fr = var == VAR_T ? s.bc_temp[id] == BC_FREE : s.bc_vel[id] == BC_FREE;
                        if (!fr) gb[var * N + id] += X[t];
                    }
                }
            }
        }
        for (size_t i = 0; i < N; i++)
            if (in_node(i))
                for (int f = 0; f < 4; f++) tb[f * N + i] *= alpha;
        const size_t st[3] = {(size_t)s.ny * s.nz, (size_t)s.nz, 1};
        const T hh[3] = {s.dx, s.dy, s.dz};
        const size_t sd = st[dir];
        const T h = hh[dir];
        const int dv = dir;  // velocity component along the sweep
        for (const Seg& sg : list)
            for (int t = 1; t < sg.size - 1; t++) {
                int i, j, k;
                s.seg_cell(sg, dir, t, i, j, k);
                const size_t p = s.idx(i, j, k);
                T lm[4], xm[4], x0[4], xp[4];
                for (int f = 0; f < 4; f++) {
                    lm[f] = lam[f * N + p];
                    xm[f] = outL[f * N + p - sd];
                    x0[f] = outL[f * N + p];
                    xp[f] = outL[f * N + p + sd];
                    inb[f * N + p] = lm[f] * 3 / dt;
                }
                // matrix adjoints per matrix (velocity fields share one)
                T abv = 0, bbv = 0, cbv = 0;
                for (int f = 0; f < 3; f++) {
                    abv -= lm[f] * xm[f];
                    bbv -= lm[f] * x0[f];
                    cbv -= lm[f] * xp[f];
                }
                T abt = -lm[3] * xm[3], bbt = -lm[3] * x0[3], cbt = -lm[3] * xp[3];
                tb[dv * N + p] += (-abv + cbv - abt + cbt) / (2 * h);
                pb[1] += (double)((-abv - cbv + 2 * bbv) / (h * h));
                pb[2] += (double)((-abt - cbt + 2 * bbt) / (h * h));
                // pressure term of the velocity along the sweep
                const T* tT = tinL + 3 * N;
                const T gT = (tT[p + sd] - tT[p - sd]) / (2 * h);
                pb[0] += (double)(-lm[dv] * gT);
                tb[3 * N + p + sd] += -s.v_T * lm[dv] / (2 * h);
                tb[3 * N + p - sd] += s.v_T * lm[dv] / (2 * h);
                // dissipation of the T row
                const T* tu = tinL;
                const T* tv = tinL + N;
                const T* tw = tinL + 2 * N;
                auto D = [&](const T* q, int ax) {
                    return (q[p + st[ax]] - q[p - st[ax]]) / (2 * hh[ax]);
                };
                auto put = [&](int f, int ax, T val) {
                    tb[f * N + p + st[ax]] += val / (2 * hh[ax]);
                    tb[f * N + p - st[ax]] -= val / (2 * hh[ax]);
                };
                const T kk = s.t_phi * lm[3];
                T phi;
                if (dir == AX_X) {
                    T u_x = D(tu, 0), v_x = D(tv, 0), w_x = D(tw, 0), u_y = D(tu, 1), u_z = D(tu, 2);
                    phi = 2 * u_x * u_x + v_x * v_x + w_x * w_x + v_x * u_y + w_x * u_z;
                    put(0, 0, kk * 4 * u_x);
                    put(1, 0, kk * (2 * v_x + u_y));
                    put(2, 0, kk * (2 * w_x + u_z));
                    put(0, 1, kk * v_x);
                    put(0, 2, kk * w_x);
                } else if (dir == AX_Y) {
                    T u_y = D(tu, 1), v_y = D(tv, 1), w_y = D(tw, 1), v_x = D(tv, 0), v_z = D(tv, 2);
                    phi = u_y * u_y + 2 * v_y * v_y + w_y * w_y + u_y * v_x + w_y * v_z;
                    put(0, 1, kk * (2 * u_y + v_x));
                    put(1, 1, kk * 4 * v_y);
                    put(2, 1, kk * (2 * w_y + v_z));
                    put(1, 0, kk * u_y);
                    put(1, 2, kk * w_y);
                } else {
                    T u_z = D(tu, 2), v_z = D(tv, 2), w_z = D(tw, 2), w_x = D(tw, 0), w_y = D(tw, 1);
                    phi = u_z * u_z + v_z * v_z + 2 * w_z * w_z + u_z * w_x + v_z * w_y;
                    put(0, 2, kk * (2 * u_z + w_x));
                    put(1, 2, kk * (2 * v_z + w_y));
                    put(2, 2, kk * 4 * w_z);
                    put(2, 0, kk * u_z);
                    put(2, 1, kk * v_z);
                }
                pb[3] += (double)(lm[3] * phi);
            }
    }

    // Reverses one recorded step: cb holds the adjoint of the step's result on
    // NODE_IN nodes and receives the adjoint of its start state (NODE_IN).
    void adj_step(const Tape& tp, std::vector<T>& cb, std::vector<T>& gb, double* pb) {
        std::vector<T> xb(4 * N, (T)0), tb(4 * N, (T)0), cacc(4 * N, (T)0), inb(4 * N, (T)0);
        for (size_t i = 0; i < N; i++)
            if (in_node(i))
                for (int f = 0; f < 4; f++) xb[f * N + i] = cb[f * N + i];
        for (int it = G - 1; it >= 0; it--)
            for (int q = 2; q >= 0; q--) {
                const int dir = q == 2 ? AX_X : q == 1 ? AX_Y : AX_Z;
                const size_t k = (size_t)3 * it + q;
                const T alpha = q == 2 ? (T)0.25 : (T)0.5, beta = q == 2 ? (T)0.75 : (T)0.5;
                std::fill(inb.begin(), inb.end(), (T)0);
                adj_sweep(dir, &tp.tin[k * 4 * N], &tp.out[k * 4 * N], xb, tb, alpha, beta, inb, gb, pb);
                if (q == 0) {
                    for (size_t i = 0; i < 4 * N; i++) cacc[i] += inb[i];
                    std::fill(xb.begin(), xb.end(), (T)0);
                } else {
                    xb = inb;
                }
            }
        // temp <- cur at the step start; UpdateBoundaries wrote the node values
        // into the state on BOUND/VALVE nodes.
        for (size_t i = 0; i < N; i++)
            for (int f = 0; f < 4; f++) {
                if (in_node(i)) cb[f * N + i] = cacc[f * N + i] + tb[f * N + i];
                else {
                    cb[f * N + i] = 0;
                    if (s.type[i] == NODE_BOUND || s.type[i] == NODE_VALVE) gb[f * N + i] += tb[f * N + i];
                }
            }
    }

    void backward(const double* seed, double* gbar, double* cbar0, double* pbar) {
        std::vector<T> cb(4 * N, (T)0), gb(4 * N, (T)0);
        for (size_t i = 0; i < 4 * N; i++) cb[i] = in_node(i % N) ? (T)seed[i] : (T)0;
        double pb[4] = {0, 0, 0, 0};
        for (int n = (int)tape.size() - 1; n >= 0; n--) adj_step(tape[n], cb, gb, pb);
        for (size_t i = 0; i < 4 * N; i++) {
            if (gbar) gbar[i] = (double)gb[i];
            if (cbar0) cbar0[i] = (double)cb[i];
        }
        for (int q = 0; q < 4; q++) pbar[q] = pb[q];
    }
};

struct Handle {
    int dtype;
    void* s;
    void* adj = nullptr;
};

}  // namespace orc

using namespace orc;

template <class T>
static Solver<T>* S(void* h) { return (Solver<T>*)((Handle*)h)->s; }

#define DISPATCH(h, ...)                                                     \
    do {                                                                      \
        if (((Handle*)(h))->dtype == 0) { auto* s = S<float>(h); __VA_ARGS__; } \
        else { auto* s = S<double>(h); __VA_ARGS__; }                         \
    } while (0)

extern "C" {

// dtype 0 = FP32 (FTYPE float), 1 = FP64.  Grid values arrive as double and
// are stored as FTYPE like the reference's Node fields.
void* orc_create(int dtype, int nx, int ny, int nz, double dx, double dy, double dz,
                 const uint8_t* type, const uint8_t* bc_vel, const uint8_t* bc_temp,
                 const double* vx, const double* vy, const double* vz, const double* T,
                 double v_T, double v_vis, double t_vis, double t_phi, int merge_mode,
                 int nblockZ) {
    Handle* h = new Handle;
    h->dtype = dtype;
    if (dtype == 0) {
        auto* s = new Solver<float>;
        s->init(nx, ny, nz, dx, dy, dz, type, bc_vel, bc_temp, vx, vy, vz, T);
        s->v_T = (float)v_T; s->v_vis = (float)v_vis; s->t_vis = (float)t_vis; s->t_phi = (float)t_phi;
        s->merge_mode = merge_mode; s->nblockZ = nblockZ;
        h->s = s;
    } else {
        auto* s = new Solver<double>;
        s->init(nx, ny, nz, dx, dy, dz, type, bc_vel, bc_temp, vx, vy, vz, T);
        s->v_T = v_T; s->v_vis = v_vis; s->t_vis = t_vis; s->t_phi = t_phi;
        s->merge_mode = merge_mode; s->nblockZ = nblockZ;
        h->s = s;
    }
    return h;
}

void orc_destroy(void* h) {
    if (((Handle*)h)->adj) {
        if (((Handle*)h)->dtype == 0) delete (Adjoint<float>*)((Handle*)h)->adj;
        else delete (Adjoint<double>*)((Handle*)h)->adj;
    }
    DISPATCH(h, delete s);
    delete (Handle*)h;
}

void orc_create_segments(void* h) { DISPATCH(h, s->create_segments()); }

int orc_num_segments(void* h, int axis) {
    int n = 0;
    DISPATCH(h, n = (int)s->segs[axis].size());
    return n;
}

// out: 7 int32 per segment (posx,posy,posz,endx,endy,endz,size)
void orc_get_segments(void* h, int axis, int32_t* out) {
    DISPATCH(h, {
        for (size_t i = 0; i < s->segs[axis].size(); i++) {
            const Seg& g = s->segs[axis][i];
            int32_t* o = out + 7 * i;
            o[0] = g.px; o[1] = g.py; o[2] = g.pz; o[3] = g.ex; o[4] = g.ey; o[5] = g.ez; o[6] = g.size;
        }
    });
}

void orc_update_boundaries(void* h) { DISPATCH(h, s->update_boundaries()); }

int orc_step(void* h, double dt, int G, int L, int compute_err, double* err) {
    int r = 0;
    DISPATCH(h, r = s->time_step((decltype(s->dx))dt, G, L, compute_err != 0, err));
    return r;
}

// which: 0 = cur, 1 = next, 2 = half, 3 = temp
void orc_get_layer_raw(void* h, int which, double* u, double* v, double* w, double* T) {
    DISPATCH(h, {
        auto* L = which == 0 ? s->cur : which == 1 ? s->next : which == 2 ? s->half : s->temp;
        double* o[4] = {u, v, w, T};
        for (int c = 0; c < 4; c++)
            for (size_t i = 0; i < s->N; i++) o[c][i] = (double)L->f[c][L->halo + i];
    });
}

void orc_set_fields(void* h, const double* u, const double* v, const double* w, const double* T) {
    DISPATCH(h, {
        const double* in[4] = {u, v, w, T};
        for (int c = 0; c < 4; c++)
            for (size_t i = 0; i < s->N; i++) s->cur->f[c][s->cur->halo + i] = (decltype(s->dx))in[c][i];
        s->copy_state_to_all();
    });
}

// Grid3D::SetData (Grid3D.h:121): new BC kinds and node values, read by the
// boundary rows (and UpdateBoundaries) from the next step on.
void orc_set_boundary(void* h, const uint8_t* bv, const uint8_t* bt, const double* vx,
                      const double* vy, const double* vz, const double* T) {
    DISPATCH(h, {
        s->bc_vel.assign(bv, bv + s->N);
        s->bc_temp.assign(bt, bt + s->N);
        for (size_t i = 0; i < s->N; i++) {
            s->gv[0][i] = (decltype(s->dx))vx[i];
            s->gv[1][i] = (decltype(s->dx))vy[i];
            s->gv[2][i] = (decltype(s->dx))vz[i];
            s->gT[i] = (decltype(s->dx))T[i];
        }
    });
}

double orc_div_error_cur(void* h) {
    double e = 0;
    DISPATCH(h, e = s->div_error(s->cur));
    return e;
}

// Solver3D::GetLayer (Solver3D.cpp:21-25): clears NODE_OUT cells of `next`
// to 99999 and resamples `next` (the buffer that holds the previous state
// after TimeStep's swap) with nearest-index filtering (TimeLayer3D.h:819-924).
void orc_get_layer(void* h, float* vel_xyz, double* outT, int ox, int oy, int oz) {
    DISPATCH(h, {
        auto* L = s->next;
        if (ox == 0) ox = s->nx;
        if (oy == 0) oy = s->ny;
        if (oz == 0) oz = s->nz;
        for (int i = 0; i < s->nx; i++)
            for (int j = 0; j < s->ny; j++)
                for (int k = 0; k < s->nz; k++)
                    if (s->type[s->idx(i, j, k)] == NODE_OUT)
                        for (int c = 0; c < 4; c++) L->e(c, i, j, k) = (decltype(s->dx))99999.0f;
        for (int i = 0; i < ox; i++)
            for (int j = 0; j < oy; j++)
                for (int k = 0; k < oz; k++) {
                    int x = i * s->nx / ox, y = j * s->ny / oy, z = k * s->nz / oz;
                    size_t ind = ((size_t)i * oy + j) * oz + k;
                    vel_xyz[3 * ind + 0] = (float)L->e(VAR_U, x, y, z);
                    vel_xyz[3 * ind + 1] = (float)L->e(VAR_V, x, y, z);
                    vel_xyz[3 * ind + 2] = (float)L->e(VAR_W, x, y, z);
                    outT[ind] = (double)L->e(VAR_T, x, y, z);
                }
    });
}

// Adjoint: `steps` recorded time steps (UpdateBoundaries + TimeStep with
// num_local = 1, CPU merge order), then the reverse pass.  seed: adjoint of
// the final state [4][N] (u, v, w, T); gbar: gradient of the node values
// [4][N]; cbar0: gradient of the initial state [4][N] (NODE_IN nodes); pbar:
// gradients of (v_T, v_vis, t_vis, t_phi).
void orc_adj_record(void* h, int steps, double dt, int G) {
    Handle* H = (Handle*)h;
    DISPATCH(h, {
        using TT = decltype(s->dx);
        if (!H->adj) H->adj = new Adjoint<TT>(*s);
        auto* A = (Adjoint<TT>*)H->adj;
        for (int i = 0; i < steps; i++) {
            s->update_boundaries();
            A->record_step((TT)dt, G);
        }
    });
}

void orc_adj_backward(void* h, const double* seed, double* gbar, double* cbar0, double* pbar) {
    Handle* H = (Handle*)h;
    DISPATCH(h, {
        using TT = decltype(s->dx);
        auto* A = (Adjoint<TT>*)H->adj;
        if (A) A->backward(seed, gbar, cbar0, pbar);
    });
}

// Standalone Thomas (Algorithms.h:21-38) for closed-form checks.  Overwrites c, d.
void orc_thomas_f64(double* a, double* b, double* c, double* d, double* x, int n) {
    solve_tridiagonal<double>(a, b, c, d, x, n);
}
void orc_thomas_f32(float* a, float* b, float* c, float* d, float* x, int n) {
    solve_tridiagonal<float>(a, b, c, d, x, n);
}

// Segment builder alone (no solver), for brute-force comparisons.
int orc_segments_count(const uint8_t* type, int nx, int ny, int nz, int axis, int nblockZ) {
    std::vector<Seg> v;
    gen_segments(type, nx, ny, nz, axis, nblockZ, v);
    return (int)v.size();
}
void orc_segments(const uint8_t* type, int nx, int ny, int nz, int axis, int nblockZ, int32_t* out) {
    std::vector<Seg> v;
    gen_segments(type, nx, ny, nz, axis, nblockZ, v);
    for (size_t i = 0; i < v.size(); i++) {
        int32_t* o = out + 7 * i;
        o[0] = v[i].px; o[1] = v[i].py; o[2] = v[i].pz; o[3] = v[i].ex; o[4] = v[i].ey; o[5] = v[i].ez; o[6] = v[i].size;
    }
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
