// C ABI and host orchestration of the ADI time step.
//
// One solver object owns the device copy of the domain (mask, boundary kinds,
// node values), five field layers and the segment lists.  Layer roles:
//   cur   state at the start of the step (TimeLayer3D `cur`)
//   next  output of the Z and X sweeps (`next`)
//   half  output of the Y sweep (`half`)
//   tA,tB the Picard layer `temp`, double-buffered: every sweep reads one buffer
//         and writes the merged values of its NODE_IN nodes into the other, so
//         the four fields of a sweep can be solved in one launch while the
//         cross-line stencils still see the pre-sweep layer (the reference's CPU
//         order, AdiSolver3D.cpp:591-604 then 640-653).
// TimeStep's extra merge after each X sweep (AdiSolver3D.cpp:347-357) is folded
// into the X sweep of the last local iteration (merge applied twice).
#include "cmcx_internal.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <cstring>
#include <memory>
#include <stdexcept>

namespace cmcx {

thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// Autotuned strided-tile variants per (grid shape, field size), process-wide.
std::map<std::string, std::array<int, 3>>& tune_cache() {
    static std::map<std::string, std::array<int, 3>> c;
    return c;
}

struct ISolver {
    virtual ~ISolver() {}
    virtual int set_options(int path, int merge_order) = 0;
    virtual int set_nblockz(int nb) = 0;
    virtual int build_segments() = 0;
    virtual int num_segments(int axis, int64_t* n) = 0;
    virtual int get_segments(int axis, int32_t* out, int64_t cap) = 0;
    virtual int get_segment_nodes(int axis, uint8_t* flags, double* vals, int64_t cap) = 0;
    virtual int update_boundaries() = 0;
    virtual int step(double dt, int G, int L, int compute_err, double* err) = 0;
    virtual int run(double dt, int steps, int G, int L, double* err) = 0;
    virtual int div_error(double* err) = 0;
    virtual int get_fields(void* u, void* v, void* w, void* p, void* T, bool device) = 0;
    virtual int set_fields(const void* u, const void* v, const void* w, const void* T) = 0;
    virtual int state_ptrs(void** p4) = 0;
    virtual int get_layer(void* vel, double* T, int ox, int oy, int oz) = 0;
    virtual int set_boundary(const uint8_t* bv, const uint8_t* bt, const double* vx,
                             const double* vy, const double* vz, const double* T) = 0;
    virtual int profile(int enable) = 0;
    virtual int profile_read(double* ms, int64_t* launches, int reset) = 0;
    virtual int step_bytes(int G, int L, double* bytes) = 0;
    virtual int autotune(int reps, int* choice) = 0;
    virtual int set_tile_variant(int axis, int variant) = 0;
    virtual int set_graph(int enable) = 0;
    virtual int batch_size() const = 0;
    virtual int select(int index) = 0;
    virtual int grid_errors(double* errs) = 0;
    virtual int set_params(const cmcx_params* p) = 0;
    virtual int adjoint_forward(double dt, int steps, int G, int every) = 0;
    virtual int adjoint_backward(const void* const* seed, void* const* gbar, void* const* sbar,
                                 double* pbar, bool device) = 0;
    virtual int adjoint_release() = 0;
    virtual int adjoint_generation(uint64_t* gen) = 0;
    virtual int synchronize() = 0;
};

// One grid of a solver: its geometry, parameters, place in the flat arrays and
// segment lists.  A solver advances all its grids together (a batch); the
// single-domain interface is a batch of one.
struct Grid {
    Geom g{};
    double h[3] = {0, 0, 0};
    double base_t = 0;
    cmcx_params prm{};
    int64_t off = 0;          // flat index of node (0, 0, 0)
    SegList segs[3];
    int free_shared = 0;      // axes with a BC_FREE node shared by two segments
    double err = 0.0;         // divergence error of the last evaluation
};

template <class T>
Layer4<T> shifted(const Layer4<T>& L, int64_t off) {
    Layer4<T> r;
    for (int f = 0; f < 4; f++) r.f[f] = L.f[f] ? L.f[f] + off : nullptr;
    return r;
}

template <class T>
class Solver final : public ISolver {
   public:
    std::vector<Grid> grids;
    int sel = 0;              // grid addressed by the per-grid calls
    int64_t NT = 0;           // length of the flat arrays (grids + alignment padding)
    hipStream_t st = nullptr;
    int path = CMCX_PATH_PLAIN;
    int merge_order = CMCX_MERGE_CPU;
    int nblockz = 1;
    bool built = false;
    double diff_error = 0.0;
    bool div_blown = false;  // the last evaluation found a grid with a non-finite error

    std::vector<uint8_t> h_type, h_bcv, h_bct;  // host copies of the flat mask / BC kinds
    uint8_t *d_type = nullptr, *d_bcv = nullptr, *d_bct = nullptr;
    uint8_t* d_tflags = nullptr;  // node flags of the fused kernels (build_node_flags)
    // the state holds the node values on BOUND/VALVE nodes with no-slip BCs:
    // set by UpdateBoundaries (and at creation, the state being the node values),
    // kept by steps, cleared when the caller replaces the state or node values
    bool bv_synced = true;
    T* d_g[4] = {nullptr, nullptr, nullptr, nullptr};
    T* buf = nullptr;  // 5 layers x 4 fields x NT
    Layer4<T> cur{}, next{}, half{}, tA{}, tB{};
    T* scratch = nullptr;
    size_t scratch_elems = 0;
    FusedWork work[3][kFusedVariants];
    InstDesc<T>* d_desc = nullptr;
    int variant[3] = {0, 0, 0};  // fused kernel variant per axis (autotuned)
    double* d_acc = nullptr;     // 2 doubles per grid
    double* d_part = nullptr;
    // node lists (fixed mask): non-IN nodes, BOUND/VALVE nodes (flat indices)
    uint32_t* d_nonin = nullptr;
    uint32_t* d_bvlist = nullptr;
    int64_t n_nonin = 0, n_bv = 0;
    // BOUND/VALVE nodes with a free BC (velocity or T): while the state is in
    // sync with the node values, the only non-IN nodes whose values move from
    // step to step (capacity n_bv; rebuilt when the BC kinds change)
    uint32_t* d_bvfree = nullptr;
    int64_t n_bvfree = 0;
    // next, tA and tB hold the state's values on every BOUND/VALVE node outside
    // d_bvfree: the step set-up need only visit d_bvfree.  True at creation
    // (every layer a copy of the state); a step in sync with the node values
    // keeps it (its no-slip rows write the value they read from the Picard
    // layer, itself the state's); a boundary fill out of sync, a step out of
    // sync, new BC kinds or a state handed out by pointer clear it.
    bool layers_ok = true;
    // GetLayer's NODE_OUT fill (Solver3D.cpp:21-25) writes 99999 into `next`,
    // the buffer that holds the previous state, and there it stays: the
    // reference's sweeps never write NODE_OUT nodes, so each of its two state
    // buffers keeps its own NODE_OUT values.  The fused Z sweep stores whole
    // runs and carries its input's NODE_OUT values into its output, which is
    // the same thing while both buffers agree.  Per grid, bit b: state buffer b
    // (b = 0 at buf, 1 after it) holds the fill.  A step between buffers that
    // disagree keeps the output buffer's own values (saved in out_keep) or
    // writes the fill again.
    std::vector<uint8_t> out_fill;
    T* keep_buf = nullptr;
    Layer4<T> out_keep{};
    // HIP graphs of whole steps, per (boundary fill, dt, G, L, state buffer)
    bool graph_mode = false;
    // (with the boundary-sync mode the sweeps were captured with)
    std::map<std::tuple<int, double, int, int, const void*, int>, hipGraphExec_t> graphs;
    hipStream_t cap = nullptr;
    // reverse mode over a recorded rollout
    struct Adj {
        bool recorded = false;
        // counts recordings and invalidations: a gradient request names the
        // recording it belongs to (cmcx_adjoint_generation)
        uint64_t generation = 0;
        std::string invalid_why;  // what cleared the last recording
        T dt = 0;
        int G = 0, steps = 0, every = 0;
        T* ck = nullptr;       // checkpoints of the state, [nck][4][NT]
        int nck = 0;
        T* tape = nullptr;     // per step: S, Picard layers 1..3G, sweep outputs 0..3G-1
        int64_t tape_layers = 0;
        int tape_G = 0;
        int nbuf = 0;              // tape buffers (two when the rollout has several intervals)
        int holds[2] = {-1, -1};   // interval each tape buffer holds (-1 = none)
        hipStream_t fs = nullptr;  // stream of the recompute, beside the reverse on `st`
        T* work = nullptr;     // xbar, tbar, state-adjoint sum, lam, node-value gradient, final state
        double* pp = nullptr;  // per-block parameter partials
        uint16_t* nmask = nullptr;  // node words of the pointwise kernel
        int32_t* tiles[3] = {nullptr, nullptr, nullptr};
        int64_t ntiles[3] = {0, 0, 0};
        int ltiles[3] = {0, 0, 0};
    } adj;
    // profiling
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
    size_t ev_used = 0;
    double prof_ms = 0.0;
    int64_t prof_launches = 0;

    Solver(const cmcx_domain* doms, const cmcx_params* ps, int count, hipStream_t stream) {
        st = stream;
        grids.resize(count);
        out_fill.assign(count, 0);
        int64_t off = 0;
        for (int i = 0; i < count; i++) {
            const cmcx_domain* d = &doms[i];
            Grid& G = grids[i];
            G.g.nx = d->dims[0];
            G.g.ny = d->dims[1];
            G.g.nz = d->dims[2];
            G.g.sy = G.g.nz;
            G.g.sx = (int64_t)G.g.ny * G.g.nz;
            G.g.N = G.g.sx * G.g.nx;
            for (int a = 0; a < 3; a++) G.h[a] = d->h[a];
            G.base_t = d->base_t;
            G.prm = ps[i];
            G.off = off;
            off += (G.g.N + 63) / 64 * 64;  // 256-byte aligned grids
        }
        NT = off;
        if (NT > ((int64_t)1 << 31) - 1) throw std::string("grids too large (total node count must fit in int32)");
        h_type.assign(NT, (uint8_t)NODE_OUT);  // padding between grids: never read
        h_bcv.assign(NT, 0);
        h_bct.assign(NT, 0);
        std::vector<T> hg[4];
        for (int f = 0; f < 4; f++) hg[f].assign(NT, (T)0);
        for (int i = 0; i < count; i++) {
            const cmcx_domain* d = &doms[i];
            const int64_t o = grids[i].off, N = grids[i].g.N;
            std::memcpy(&h_type[o], d->type, N);
            std::memcpy(&h_bcv[o], d->bc_vel, N);
            std::memcpy(&h_bct[o], d->bc_temp, N);
            const double* src[4] = {d->v[0], d->v[1], d->v[2], d->T};
            for (int f = 0; f < 4; f++)
                for (int64_t k = 0; k < N; k++) hg[f][o + k] = (T)src[f][k];
        }
        CMCX_HIP(hipMalloc((void**)&d_type, NT));
        CMCX_HIP(hipMalloc((void**)&d_bcv, NT));
        CMCX_HIP(hipMalloc((void**)&d_bct, NT));
        CMCX_HIP(hipMalloc((void**)&d_tflags, NT));
        for (int f = 0; f < 4; f++) CMCX_HIP(hipMalloc((void**)&d_g[f], sizeof(T) * NT));
        CMCX_HIP(hipMalloc((void**)&buf, sizeof(T) * NT * 20));
        CMCX_HIP(hipMemsetAsync(buf, 0, sizeof(T) * NT * 20, st));
        Layer4<T>* Ls[5] = {&cur, &next, &half, &tA, &tB};
        for (int l = 0; l < 5; l++)
            for (int f = 0; f < 4; f++) Ls[l]->f[f] = buf + (int64_t)(l * 4 + f) * NT;
        CMCX_HIP(hipMalloc((void**)&d_acc, sizeof(double) * 2 * count));
        CMCX_HIP(hipMalloc((void**)&d_part, sizeof(double) * div_partials()));
        CMCX_HIP(hipMalloc((void**)&d_desc, sizeof(InstDesc<T>) * count));
        CMCX_HIP(hipMemcpyAsync(d_type, h_type.data(), NT, hipMemcpyHostToDevice, st));
        CMCX_HIP(hipMemcpyAsync(d_bcv, h_bcv.data(), NT, hipMemcpyHostToDevice, st));
        CMCX_HIP(hipMemcpyAsync(d_bct, h_bct.data(), NT, hipMemcpyHostToDevice, st));
        upload_flags(0, NT);
        for (int f = 0; f < 4; f++)
            CMCX_HIP(hipMemcpyAsync(d_g[f], hg[f].data(), sizeof(T) * NT, hipMemcpyHostToDevice, st));
        {
            std::vector<uint32_t> nonin, bvl;
            for (const Grid& G : grids)
                for (int64_t k = 0; k < G.g.N; k++) {
                    const int64_t i = G.off + k;
                    const uint8_t t = h_type[i];
                    if (t != NODE_IN) nonin.push_back((uint32_t)i);
                    if (t == NODE_BOUND || t == NODE_VALVE) bvl.push_back((uint32_t)i);
                }
            n_nonin = (int64_t)nonin.size();
            n_bv = (int64_t)bvl.size();
            CMCX_HIP(hipMalloc((void**)&d_nonin, sizeof(uint32_t) * std::max<int64_t>(n_nonin, 1)));
            CMCX_HIP(hipMalloc((void**)&d_bvlist, sizeof(uint32_t) * std::max<int64_t>(n_bv, 1)));
            CMCX_HIP(hipMalloc((void**)&d_bvfree, sizeof(uint32_t) * std::max<int64_t>(n_bv, 1)));
            if (n_nonin)
                CMCX_HIP(hipMemcpyAsync(d_nonin, nonin.data(), sizeof(uint32_t) * n_nonin,
                                        hipMemcpyHostToDevice, st));
            if (n_bv)
                CMCX_HIP(hipMemcpyAsync(d_bvlist, bvl.data(), sizeof(uint32_t) * n_bv,
                                        hipMemcpyHostToDevice, st));
            // TimeLayer3D(backend, grid) -> CopyFromGrid (TimeLayer3D.h:734-781)
            for (int f = 0; f < 4; f++)
                CMCX_HIP(hipMemcpyAsync(cur.f[f], d_g[f], sizeof(T) * NT, hipMemcpyDeviceToDevice, st));
            copy_state_to_all(0, NT);
            CMCX_HIP(hipStreamSynchronize(st));  // host vectors go out of scope
        }
        upload_free_list();
        upload_desc();
    }

    void upload_free_list() {
        std::vector<uint32_t> fl;
        for (const Grid& G : grids)
            for (int64_t k = 0; k < G.g.N; k++) {
                const int64_t i = G.off + k;
                const uint8_t t = h_type[i];
                if ((t == NODE_BOUND || t == NODE_VALVE) && (h_bcv[i] == BC_FREE || h_bct[i] == BC_FREE))
                    fl.push_back((uint32_t)i);
            }
        n_bvfree = (int64_t)fl.size();
        if (n_bvfree)
            CMCX_HIP(hipMemcpyAsync(d_bvfree, fl.data(), sizeof(uint32_t) * n_bvfree,
                                    hipMemcpyHostToDevice, st));
        CMCX_HIP(hipStreamSynchronize(st));  // fl goes out of scope
    }

    void upload_flags(int64_t off, int64_t n) {
        std::vector<uint8_t> h(n);
        build_node_flags(&h_type[off], &h_bcv[off], &h_bct[off], n, h.data());
        CMCX_HIP(hipMemcpyAsync(d_tflags + off, h.data(), n, hipMemcpyHostToDevice, st));
        CMCX_HIP(hipStreamSynchronize(st));  // h goes out of scope
    }

    // The reference allocates next, half and temp without initialising them
    // (TimeLayer3D.h:1065-1075); here every layer starts as a copy of the state,
    // and so after set_fields.  Nodes that no segment covers then hold the same
    // value in cur and next, which lets the fused Z sweep write its input there
    // (a whole-vector store) without changing any result.
    void copy_state_to_all(int64_t off, int64_t n) {
        const Layer4<T> src = shifted(cur, off);
        for (const Layer4<T>* L : {&next, &half, &tA, &tB}) launch_copy4<T>(shifted(*L, off), src, n, st);
    }

    // Descriptor table of the fused kernels (dt-independent part of SweepParams).
    void upload_desc() {
        std::vector<InstDesc<T>> h(grids.size());
        for (size_t i = 0; i < grids.size(); i++) {
            h[i].g = grids[i].g;
            h[i].off = grids[i].off;
            for (int a = 0; a < 3; a++) h[i].p[a] = sweep_params((int)i, a, (T)1);
        }
        CMCX_HIP(hipMemcpyAsync(d_desc, h.data(), sizeof(InstDesc<T>) * h.size(), hipMemcpyHostToDevice, st));
        CMCX_HIP(hipStreamSynchronize(st));
    }

    ~Solver() override {
        hipStreamSynchronize(st);
        clear_graphs();
        if (cap) hipStreamDestroy(cap);
        hipFree(d_type);
        hipFree(d_bcv);
        hipFree(d_bct);
        hipFree(d_tflags);
        for (auto* p : d_g) hipFree(p);
        hipFree(buf);
        hipFree(keep_buf);
        for (auto& G : grids)
            for (auto& s : G.segs) hipFree(s.d);
        hipFree(scratch);
        for (auto& wa : work)
            for (auto& w : wa) w.release();
        hipFree(d_desc);
        hipFree(d_acc);
        hipFree(d_part);
        hipFree(d_nonin);
        hipFree(d_bvlist);
        hipFree(d_bvfree);
        hipFree(adj.ck);
        hipFree(adj.tape);
        if (adj.fs) hipStreamDestroy(adj.fs);
        hipFree(adj.work);
        hipFree(adj.pp);
        hipFree(adj.nmask);
        for (auto* t : adj.tiles) hipFree(t);
        for (auto& e : ev_pool) {
            hipEventDestroy(e.first);
            hipEventDestroy(e.second);
        }
    }

    Grid& G() { return grids[sel]; }

    int check_mask() {
        unsigned long long* d_out = nullptr;
        unsigned long long h[2] = {0, 0};
        CMCX_HIP(hipMalloc((void**)&d_out, sizeof(h)));
        CMCX_HIP(hipMemsetAsync(d_out, 0, sizeof(h), st));
        for (const Grid& Gr : grids) launch_mask_check(Gr.g, d_type + Gr.off, d_out, st);
        CMCX_HIP(hipMemcpyAsync(h, d_out, sizeof(h), hipMemcpyDeviceToHost, st));
        CMCX_HIP(hipStreamSynchronize(st));
        CMCX_HIP(hipFree(d_out));
        if (h[0] || h[1]) {
            set_error("mask not well formed: " + std::to_string(h[0]) +
                      " NODE_IN nodes on a box face, " + std::to_string(h[1]) +
                      " NODE_OUT nodes next to NODE_IN");
            return CMCX_ERR_MASK;
        }
        return CMCX_OK;
    }

    void clear_graphs() {
        for (auto& kv : graphs) hipGraphExecDestroy(kv.second);
        graphs.clear();
    }

    int set_options(int p, int m) override {
        if (p != CMCX_PATH_PLAIN && p != CMCX_PATH_FUSED) return CMCX_ERR_ARG;
        if (m != CMCX_MERGE_CPU && m != CMCX_MERGE_GPU_ORDER) return CMCX_ERR_ARG;
        if (p == CMCX_PATH_FUSED && m != CMCX_MERGE_CPU) {
            set_error("the fused path implements the CPU merge order only");
            return CMCX_ERR_UNSUPPORTED;
        }
        if (p == CMCX_PATH_FUSED)
            for (const Grid& Gr : grids)
                if (!fused_supported<T>(Gr.g)) {
                    set_error("fused path: grid lines longer than supported");
                    return CMCX_ERR_UNSUPPORTED;
                }
        if (p != path || m != merge_order) invalidate_recording("set_options");
        path = p;
        merge_order = m;
        clear_graphs();
        if (built) prepare_path_buffers();
        return CMCX_OK;
    }

    int set_nblockz(int nb) override {
        for (const Grid& Gr : grids)
            if (nb < 1 || nb > Gr.g.nz || nb > 64) return CMCX_ERR_ARG;
        nblockz = nb;
        built = false;
        clear_graphs();
        return CMCX_OK;
    }

    // The per-segment kernel runs an axis on the plain path, and for every axis
    // on which some grid has a BC_FREE node shared by two segments.
    bool plain_axis(int a) const {
        if (path == CMCX_PATH_PLAIN) return true;
        for (const Grid& Gr : grids)
            if ((Gr.free_shared >> a) & 1) return true;
        return false;
    }

    void prepare_path_buffers() {
        size_t need = 0;
        for (int a = 0; a < 3; a++)
            if (plain_axis(a))
                for (const Grid& Gr : grids)
                    need = std::max(need, (size_t)Gr.segs[a].max_size * 4 * (size_t)Gr.segs[a].n * 2);
        if (need > scratch_elems) {
            hipFree(scratch);
            scratch = nullptr;
            CMCX_HIP(hipMalloc((void**)&scratch, sizeof(T) * need));
            scratch_elems = need;
        }
        if (path == CMCX_PATH_FUSED)
            for (int a = 0; a < 3; a++)
                if (!plain_axis(a)) ensure_work(a, a == AX_Z ? 0 : variant[a]);
        clear_graphs();
    }

    void ensure_work(int a, int v) {
        if (work[a][v].maxn) return;  // built
        std::vector<Geom> gs;
        std::vector<int64_t> offs;
        for (const Grid& Gr : grids) {
            gs.push_back(Gr.g);
            offs.push_back(Gr.off);
        }
        build_fused_work(h_type.data(), gs, offs, a, strided_lpg<T>(v), strided_chunk_rows<T>(v),
                         sizeof(T), work[a][v]);
    }

    int build_segments() override {
        for (Grid& Gr : grids) {
            for (int a = 0; a < 3; a++) build_segment_list(d_type + Gr.off, Gr.g, a, nblockz, Gr.segs[a], st);
            Gr.free_shared = shared_free_axes(&h_type[Gr.off], &h_bcv[Gr.off], &h_bct[Gr.off], Gr.g);
        }
        auto it = tune_cache().find(tune_key());
        if (it != tune_cache().end())
            for (int a = 0; a < 3; a++) variant[a] = it->second[a];
        prepare_path_buffers();
        built = true;
        return CMCX_OK;
    }

    int num_segments(int axis, int64_t* n) override {
        if (axis < 0 || axis > 2 || !n) return CMCX_ERR_ARG;
        if (!built) return CMCX_ERR_STATE;
        *n = G().segs[axis].n;
        return CMCX_OK;
    }

    int get_segments(int axis, int32_t* out, int64_t cap) override {
        if (axis < 0 || axis > 2 || !out) return CMCX_ERR_ARG;
        if (!built) return CMCX_ERR_STATE;
        const SegList& s = G().segs[axis];
        if (cap < s.n) return CMCX_ERR_ARG;
        if (s.n == 0) return CMCX_OK;
        std::vector<int32_t> h(7 * s.n);
        CMCX_HIP(hipMemcpyAsync(h.data(), s.d, sizeof(int32_t) * 7 * s.n, hipMemcpyDeviceToHost, st));
        CMCX_HIP(hipStreamSynchronize(st));
        for (int64_t i = 0; i < s.n; i++)
            for (int c = 0; c < 7; c++) out[7 * i + c] = h[c * s.n + i];
        return CMCX_OK;
    }

    int get_segment_nodes(int axis, uint8_t* flags, double* vals, int64_t cap) override {
        if (axis < 0 || axis > 2 || !flags || !vals) return CMCX_ERR_ARG;
        if (!built) return CMCX_ERR_STATE;
        const Grid& Gr = G();
        const Geom& g = Gr.g;
        int64_t n = Gr.segs[axis].n;
        if (cap < n) return CMCX_ERR_ARG;
        std::vector<int32_t> sl(7 * std::max<int64_t>(n, 1));
        if (n) {
            std::vector<int32_t> h(7 * n);
            CMCX_HIP(hipMemcpyAsync(h.data(), Gr.segs[axis].d, sizeof(int32_t) * 7 * n,
                                    hipMemcpyDeviceToHost, st));
            CMCX_HIP(hipStreamSynchronize(st));
            for (int64_t i = 0; i < n; i++)
                for (int c = 0; c < 7; c++) sl[7 * i + c] = h[c * n + i];
        }
        // node records from the device copy of the grid
        const int64_t N = g.N;
        std::vector<T> gv[4];
        for (int f = 0; f < 4; f++) {
            gv[f].resize(N);
            CMCX_HIP(hipMemcpyAsync(gv[f].data(), d_g[f] + Gr.off, sizeof(T) * N,
                                    hipMemcpyDeviceToHost, st));
        }
        CMCX_HIP(hipStreamSynchronize(st));
        for (int64_t i = 0; i < n; i++) {
            const int32_t* q = &sl[7 * i];
            int64_t ids[2] = {q[0] * g.sx + q[1] * g.sy + q[2], q[3] * g.sx + q[4] * g.sy + q[5]};
            for (int e = 0; e < 2; e++) {
                int64_t id = ids[e];
                flags[6 * i + 3 * e + 0] = h_type[Gr.off + id];
                flags[6 * i + 3 * e + 1] = h_bcv[Gr.off + id];
                flags[6 * i + 3 * e + 2] = h_bct[Gr.off + id];
                for (int f = 0; f < 4; f++) vals[8 * i + 4 * e + f] = (double)gv[f][id];
            }
        }
        return CMCX_OK;
    }

    // Boundary fill (AdiSolver3D.cpp:286-295): node values into the state on
    // BOUND and VALVE nodes.  While the state is in sync, its no-slip nodes
    // already hold them and only the free ones are visited.
    void enqueue_update_boundaries(bool fast) {
        if (fast) launch_update_boundaries_list<T>(n_bvfree, d_bvfree, d_g, cur, st);
        else launch_update_boundaries_list<T>(n_bv, d_bvlist, d_g, cur, st);
    }

    // The flags after a boundary fill; returns whether the fill may be the short one.
    bool note_update_boundaries() {
        const bool fast = bv_synced;
        if (!fast) layers_ok = false;  // no-slip state values change under next, tA, tB
        bv_synced = true;
        return fast;
    }

    int update_boundaries() override {
        enqueue_update_boundaries(note_update_boundaries());
        return CMCX_OK;
    }

    SweepParams<T> sweep_params(int gi, int axis, T dt) const {
        const Grid& Gr = grids[gi];
        SweepParams<T> p;
        p.dt = dt;
        p.hx = (T)Gr.h[0];
        p.hy = (T)Gr.h[1];
        p.hz = (T)Gr.h[2];
        p.h = axis == AX_X ? p.hx : axis == AX_Y ? p.hy : p.hz;
        T v_vis = (T)Gr.prm.v_vis, t_vis = (T)Gr.prm.t_vis;
        p.vis_v = v_vis / (p.h * p.h);
        p.vis_t = t_vis / (p.h * p.h);
        p.v_T = (T)Gr.prm.v_T;
        p.t_phi = (T)Gr.prm.t_phi;
        p.three_dt = (T)3 / dt;
        p.inv_2h = (T)1 / (2 * p.h);
        p.inv_2hx = (T)1 / (2 * p.hx);
        p.inv_2hy = (T)1 / (2 * p.hy);
        p.inv_2hz = (T)1 / (2 * p.hz);
        return p;
    }

    hipEvent_t* next_events() {
        if (ev_used == ev_pool.size()) {
            hipEvent_t a, b;
            CMCX_HIP(hipEventCreate(&a));
            CMCX_HIP(hipEventCreate(&b));
            ev_pool.push_back({a, b});
        }
        return &ev_pool[ev_used++].first;
    }

    void harvest_events() {
        if (!ev_used) return;
        CMCX_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < ev_used; i++) {
            float ms = 0.f;
            CMCX_HIP(hipEventElapsedTime(&ms, ev_pool[i].first, ev_pool[i].second));
            prof_ms += ms;
        }
        prof_launches += (int64_t)ev_used;
        ev_used = 0;
    }

    // stores (fused path): bit 0 writes the output, bit 1 the merged Picard layer
    void sweep(int axis, Layer4<T> in, Layer4<T> tin, Layer4<T> out, Layer4<T> tout, int merges,
               int var0, int var1, T dt, int stores = 3) {
        hipEvent_t* ev = nullptr;
        if (prof) {
            if (ev_used >= 4096) harvest_events();
            ev = next_events();
            CMCX_HIP(hipEventRecord(ev[0], st));
        }
        if (plain_axis(axis)) {
            for (int i = 0; i < (int)grids.size(); i++) {
                const Grid& Gr = grids[i];
                const int64_t o = Gr.off;
                const T* gs[4] = {d_g[0] + o, d_g[1] + o, d_g[2] + o, d_g[3] + o};
                launch_plain_sweep<T>(Gr.g, axis, Gr.segs[axis], sweep_params(i, axis, dt), d_type + o,
                                      d_bcv + o, d_bct + o, gs, shifted(in, o), shifted(tin, o),
                                      shifted(out, o), shifted(tout, o), merges, var0, var1, scratch, st);
            }
        } else {
            const int v = variant[axis];
            const int lv = axis == AX_Z ? 0 : v;
            launch_fused_sweep<T>(d_desc, (T)3 / dt, axis, work[axis][lv], d_tflags, d_g, in, tin,
                                  out, tout, merges, bv_synced ? 1 : 0, stores, v, st);
        }
        if (prof) CMCX_HIP(hipEventRecord(ev[1], st));
    }

    // Enqueues the sweeps of one time step (no synchronisation, no host reads);
    // `short_prep`: the step set-up visits only the free BOUND/VALVE nodes.
    void enqueue_step(T dt, int G, int L, bool short_prep) {
        if (merge_order == CMCX_MERGE_CPU) {
            // TimeStep set-up (AdiSolver3D.cpp:308-316): next <- cur on BOUND and
            // VALVE nodes, temp <- cur (the sweeps write NODE_IN nodes of temp)
            if (short_prep) launch_prep_step_list<T>(n_bvfree, d_bvfree, d_type, cur, next, tA, tB, st);
            else launch_prep_step_list<T>(n_nonin, d_nonin, d_type, cur, next, tA, tB, st);
            Layer4<T> tin = cur;  // temp == cur at the first sweep
            auto other = [&](const Layer4<T>& t) { return t.f[0] == tB.f[0] ? tA : tB; };
            for (int it = 0; it < G; it++) {
                const int axes[3] = {AX_Z, AX_Y, AX_X};
                Layer4<T> ins[3] = {cur, next, half};
                Layer4<T> outs[3] = {next, half, next};
                for (int q = 0; q < 3; q++) {
                    int a = axes[q];
                    for (int l = 0; l < L; l++) {
                        Layer4<T> tout = other(tin);
                        int merges = (a == AX_X && l == L - 1) ? 2 : 1;
                        // Dead stores the fused sweeps skip (the results do not
                        // change): an output that the next local sweep overwrites,
                        // the X output of every Picard iteration but the last (the
                        // next iteration's Z sweep rewrites it before anything
                        // reads it; the Y sweep reads it on NODE_IN nodes only,
                        // which the Z sweep writes), and the Picard layer after
                        // the step's last sweep (the next step starts from the state).
                        int stores = 3;
                        if (l < L - 1 || (a == AX_X && it < G - 1)) stores &= ~1;
                        if (a == AX_X && it == G - 1 && l == L - 1) stores &= ~2;
                        sweep(a, ins[q], tin, outs[q], tout, merges, 0, 4, dt, stores);
                        tin = tout;
                    }
                }
            }
        } else {
            // the reference's CUDA order: temp <- cur, then per field in-place merges
            Layer4<T> none{};
            Geom flat{};
            flat.N = NT;
            launch_prep_step<T>(flat, d_type, cur, next, tA, none, st);
            for (int it = 0; it < G; it++) {
                const int axes[3] = {AX_Z, AX_Y, AX_X};
                Layer4<T> ins[3] = {cur, next, half};
                Layer4<T> outs[3] = {next, half, next};
                for (int q = 0; q < 3; q++)
                    for (int l = 0; l < L; l++)
                        for (int v = 0; v < 4; v++)
                            sweep(axes[q], ins[q], tA, outs[q], tA, 1, v, v + 1, dt);
            }
        }
    }

    // The graph of (boundary fill +) one step for the current state buffer;
    // captured on a private stream the first time, replayed on the solver's stream.
    // ub: 0 no fill, 1 the whole list, 2 the free nodes only; the sweeps read the
    // boundary-sync mode from bv_synced, as set for after the fill.
    hipGraphExec_t step_graph(int ub, bool short_prep, T dt, int G, int L) {
        auto key = std::make_tuple(ub, (double)dt, G, L, (const void*)cur.f[0],
                                   (short_prep ? 2 : 0) + (bv_synced ? 1 : 0));
        auto it = graphs.find(key);
        if (it != graphs.end()) return it->second;
        if (!cap) CMCX_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipStream_t user = st;
        hipGraph_t gr = nullptr;
        st = cap;
        CMCX_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        try {
            if (ub) enqueue_update_boundaries(ub == 2);
            enqueue_step(dt, G, L, short_prep);
        } catch (...) {
            hipStreamEndCapture(cap, &gr);
            if (gr) hipGraphDestroy(gr);
            st = user;
            throw;
        }
        st = user;
        CMCX_HIP(hipStreamEndCapture(cap, &gr));
        hipGraphExec_t ex = nullptr;
        hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        hipGraphDestroy(gr);
        CMCX_HIP(e);
        graphs[key] = ex;
        return ex;
    }

    // (Boundary fill +) one step, replayed as a graph or launched directly.
    void launch_step(bool ub, T dt, int G, int L) {
        const int ubm = ub ? (note_update_boundaries() ? 2 : 1) : 0;
        const bool short_prep = layers_ok;
        if (graph_mode && !prof) {
            CMCX_HIP(hipGraphLaunch(step_graph(ubm, short_prep, dt, G, L), st));
        } else {
            if (ubm) enqueue_update_boundaries(ubm == 2);
            enqueue_step(dt, G, L, short_prep);
        }
        // a step in sync writes the state's own values on no-slip nodes (the
        // CUDA merge order's set-up leaves tB alone)
        layers_ok = merge_order == CMCX_MERGE_CPU && bv_synced;
    }

    int state_buffer(const Layer4<T>& L) const { return L.f[0] == buf ? 0 : 1; }

    // launch_step, with each state buffer keeping its own NODE_OUT values
    // (see out_fill).
    void step_keeping_out_fill(bool ub, T dt, int G, int L) {
        const int pc = state_buffer(cur), pn = state_buffer(next);
        std::vector<int> differ;
        for (size_t i = 0; i < grids.size(); i++)
            if (((out_fill[i] >> pc) ^ (out_fill[i] >> pn)) & 1) differ.push_back((int)i);
        for (int i : differ) {
            const Grid& Gr = grids[i];
            if ((out_fill[i] >> pn) & 1) continue;
            if (!keep_buf) {
                CMCX_HIP(hipMalloc((void**)&keep_buf, sizeof(T) * NT * 4));
                for (int f = 0; f < 4; f++) out_keep.f[f] = keep_buf + (int64_t)f * NT;
            }
            launch_copy_typed<T>(Gr.g, d_type + Gr.off, NODE_OUT, shifted(next, Gr.off),
                                 shifted(out_keep, Gr.off), st);
        }
        launch_step(ub, dt, G, L);
        for (int i : differ) {
            const Grid& Gr = grids[i];
            if ((out_fill[i] >> pn) & 1) {
                const T miss = (T)99999.0f;
                launch_fill_typed<T>(Gr.g, d_type + Gr.off, NODE_OUT, shifted(next, Gr.off), miss,
                                     miss, miss, miss, st);
            } else {
                launch_copy_typed<T>(Gr.g, d_type + Gr.off, NODE_OUT, shifted(out_keep, Gr.off),
                                     shifted(next, Gr.off), st);
            }
        }
    }

    // Launches the step (graph or direct), then the divergence guard and swap.
    int advance(bool ub, double dt_d, int G, int L, int compute_err, double* err) {
        if (!built) {
            set_error("step before build_segments");
            return CMCX_ERR_STATE;
        }
        if (G < 0 || L < 1) return CMCX_ERR_ARG;
        const T dt = (T)dt_d;
        step_keeping_out_fill(ub, dt, G, L);
        if (compute_err) diff_error = eval_div(next);
        if (err) *err = diff_error;
        if (prof) harvest_events();
        // the reference's guard (AdiSolver3D.cpp:371-374) tests err > 0.01, which
        // a NaN error passes; here a non-finite error of a grid with fluid nodes
        // (a state that blew up) fires it too.  A grid without fluid nodes
        // divides 0 by 0 and passes, as in the reference.
        if (diff_error > 0.01 || div_blown) {
            set_error("divergence error above threshold (0.01)");
            return CMCX_DIVERGED;
        }
        std::swap(cur, next);
        return CMCX_OK;
    }

    int step(double dt, int G, int L, int compute_err, double* err) override {
        return advance(false, dt, G, L, compute_err, err);
    }

    // The reference's time loop (FluidSolver3D.cpp:226-242): boundary fill and
    // step, error every 10th step and on the last.
    int run(double dt, int steps, int G, int L, double* err) override {
        for (int i = 0; i < steps; i++) {
            int rc = advance(true, dt, G, L, (i % 10 == 0) || (i == steps - 1), err);
            if (rc != CMCX_OK) return rc;
        }
        return CMCX_OK;
    }

    // Divergence error of every grid of layer Lr (stored per grid); returns the largest.
    double eval_div(Layer4<T> Lr) {
        const int n = (int)grids.size();
        for (int i = 0; i < n; i++) {
            const Grid& Gr = grids[i];
            launch_div_error<T>(Gr.g, d_type + Gr.off, shifted(Lr, Gr.off), (T)Gr.h[0], (T)Gr.h[1],
                                (T)Gr.h[2], d_acc + 2 * i, d_part, st);
        }
        std::vector<double> h(2 * n);
        CMCX_HIP(hipMemcpyAsync(h.data(), d_acc, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, st));
        CMCX_HIP(hipStreamSynchronize(st));
        double worst = 0.0;
        div_blown = false;
        for (int i = 0; i < n; i++) {
            grids[i].err = h[2 * i] / h[2 * i + 1];
            const bool blown = h[2 * i + 1] > 0 && !std::isfinite(grids[i].err);
            div_blown = div_blown || blown;
            worst = i == 0 ? grids[i].err : std::max(worst, grids[i].err);
        }
        // a blown-up grid's NaN is the result (std::max drops it after the first)
        if (div_blown && !std::isinf(worst)) worst = std::nan("");
        return worst;
    }

    int div_error(double* err) override {
        if (!err) return CMCX_ERR_ARG;
        const bool blown = div_blown;  // the step guard's state is the last step's
        eval_div(cur);
        div_blown = blown;
        *err = G().err;
        return CMCX_OK;
    }

    int grid_errors(double* errs) override {
        if (!errs) return CMCX_ERR_ARG;
        for (size_t i = 0; i < grids.size(); i++) errs[i] = grids[i].err;
        return CMCX_OK;
    }

    int get_fields(void* u, void* v, void* w, void* p, void* Tt, bool device) override {
        const Grid& Gr = G();
        const int64_t N = Gr.g.N, o = Gr.off;
        hipMemcpyKind k = device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        void* dst[4] = {u, v, w, Tt};
        for (int f = 0; f < 4; f++)
            if (dst[f]) CMCX_HIP(hipMemcpyAsync(dst[f], cur.f[f] + o, sizeof(T) * N, k, st));
        if (p) {
            if (device) {
                launch_scale<T>(cur.f[3] + o, (T*)p, (T)Gr.prm.v_T, N, st);
            } else {
                // p = v_T * T, the ideal-gas closure the reference's momentum rows use
                std::vector<T> tmp(N);
                CMCX_HIP(hipMemcpyAsync(tmp.data(), cur.f[3] + o, sizeof(T) * N, hipMemcpyDeviceToHost, st));
                CMCX_HIP(hipStreamSynchronize(st));
                T vT = (T)Gr.prm.v_T;
                T* pp = (T*)p;
                for (int64_t i = 0; i < N; i++) pp[i] = vT * tmp[i];
            }
        }
        CMCX_HIP(hipStreamSynchronize(st));
        return CMCX_OK;
    }

    int set_fields(const void* u, const void* v, const void* w, const void* Tt) override {
        const Grid& Gr = G();
        const void* src[4] = {u, v, w, Tt};
        for (int f = 0; f < 4; f++)
            if (src[f])
                CMCX_HIP(hipMemcpyAsync(cur.f[f] + Gr.off, src[f], sizeof(T) * Gr.g.N,
                                        hipMemcpyHostToDevice, st));
        copy_state_to_all(Gr.off, Gr.g.N);
        out_fill[sel] = 0;
        bv_synced = false;
        CMCX_HIP(hipStreamSynchronize(st));
        return CMCX_OK;
    }

    int state_ptrs(void** p4) override {
        if (!p4) return CMCX_ERR_ARG;
        bv_synced = false;  // the caller may write the state
        layers_ok = false;
        for (int f = 0; f < 4; f++) p4[f] = cur.f[f] + G().off;
        return CMCX_OK;
    }

    int get_layer(void* vel, double* Tout, int ox, int oy, int oz) override {
        if (!vel || !Tout || ox < 0 || oy < 0 || oz < 0) return CMCX_ERR_ARG;
        const Grid& Gr = G();
        const Geom& g = Gr.g;
        if (ox == 0) ox = g.nx;
        if (oy == 0) oy = g.ny;
        if (oz == 0) oz = g.nz;
        // Solver3D::GetLayer: next->Clear(grid, NODE_OUT, 99999, ...) then FilterToArrays
        T miss = (T)99999.0f;
        launch_fill_typed<T>(g, d_type + Gr.off, NODE_OUT, shifted(next, Gr.off), miss, miss, miss, miss, st);
        out_fill[sel] |= (uint8_t)(1 << state_buffer(next));
        const int64_t N = g.N;
        std::vector<T> h[4];
        for (int f = 0; f < 4; f++) {
            h[f].resize(N);
            CMCX_HIP(hipMemcpyAsync(h[f].data(), next.f[f] + Gr.off, sizeof(T) * N,
                                    hipMemcpyDeviceToHost, st));
        }
        CMCX_HIP(hipStreamSynchronize(st));
        T* V = (T*)vel;
        for (int i = 0; i < ox; i++)
            for (int j = 0; j < oy; j++)
                for (int k = 0; k < oz; k++) {
                    int x = i * g.nx / ox, y = j * g.ny / oy, z = k * g.nz / oz;
                    int64_t src = x * g.sx + y * g.sy + z;
                    int64_t ind = ((int64_t)i * oy + j) * oz + k;
                    V[3 * ind + 0] = h[0][src];
                    V[3 * ind + 1] = h[1][src];
                    V[3 * ind + 2] = h[2][src];
                    Tout[ind] = (double)h[3][src];
                }
        return CMCX_OK;
    }

    int set_boundary(const uint8_t* bv, const uint8_t* bt, const double* vx, const double* vy,
                     const double* vz, const double* Tt) override {
        if (!bv || !bt || !vx || !vy || !vz || !Tt) return CMCX_ERR_ARG;
        Grid& Gr = G();
        const int64_t N = Gr.g.N, o = Gr.off;
        for (int64_t i = 0; i < N; i++)
            if (bv[i] > 1 || bt[i] > 1) {
                set_error("BC kind out of range at node " + std::to_string(i));
                return CMCX_ERR_ARG;
            }
        invalidate_recording("set_boundary");
        std::memcpy(&h_bcv[o], bv, N);
        std::memcpy(&h_bct[o], bt, N);
        CMCX_HIP(hipMemcpyAsync(d_bcv + o, bv, N, hipMemcpyHostToDevice, st));
        CMCX_HIP(hipMemcpyAsync(d_bct + o, bt, N, hipMemcpyHostToDevice, st));
        std::vector<T> tmp(N);
        const double* src[4] = {vx, vy, vz, Tt};
        for (int f = 0; f < 4; f++) {
            for (int64_t i = 0; i < N; i++) tmp[i] = (T)src[f][i];
            CMCX_HIP(hipMemcpyAsync(d_g[f] + o, tmp.data(), sizeof(T) * N, hipMemcpyHostToDevice, st));
            CMCX_HIP(hipStreamSynchronize(st));  // tmp is reused
        }
        upload_flags(o, N);
        upload_free_list();
        bv_synced = false;
        layers_ok = false;
        clear_graphs();  // the free-node count is captured
        if (built) {
            Gr.free_shared = shared_free_axes(&h_type[o], &h_bcv[o], &h_bct[o], Gr.g);
            prepare_path_buffers();
        }
        return CMCX_OK;
    }

    int profile(int enable) override {
        if (prof && !enable) harvest_events();
        prof = enable != 0;
        return CMCX_OK;
    }

    int profile_read(double* ms, int64_t* launches, int reset) override {
        harvest_events();
        if (ms) *ms = prof_ms;
        if (launches) *launches = prof_launches;
        if (reset) {
            prof_ms = 0.0;
            prof_launches = 0;
        }
        return CMCX_OK;
    }

    // Times every variant of the fused X, Y and Z sweeps on these grids (scratch
    // layers only: the state is untouched) and keeps the fastest.
    int autotune(int reps, int* choice) override {
        if (!built) return CMCX_ERR_STATE;
        if (reps < 1) reps = 3;
        if (path == CMCX_PATH_FUSED) {
            const T dt = (T)1e-4;
            hipEvent_t e0, e1;
            CMCX_HIP(hipEventCreate(&e0));
            CMCX_HIP(hipEventCreate(&e1));
            for (int a = 0; a < 3; a++) {
                if (plain_axis(a)) continue;
                float best = 1e30f;
                int bestv = variant[a];
                for (int v = 0; v < (a == AX_Z ? kZVariants : kFusedVariants); v++) {
                    if (a != AX_Z && !variant_fits(a, v)) continue;
                    const int lv = a == AX_Z ? 0 : v;  // the Z work list does not depend on v
                    ensure_work(a, lv);
                    auto launch = [&]() {
                        launch_fused_sweep<T>(d_desc, (T)3 / dt, a, work[a][lv], d_tflags, d_g,
                                              cur, cur, half, tA, 1, bv_synced ? 1 : 0, 3, v, st);
                    };
                    launch();  // warm-up
                    CMCX_HIP(hipEventRecord(e0, st));
                    for (int r = 0; r < reps; r++) launch();
                    CMCX_HIP(hipEventRecord(e1, st));
                    CMCX_HIP(hipEventSynchronize(e1));
                    float ms = 0.f;
                    CMCX_HIP(hipEventElapsedTime(&ms, e0, e1));
                    if (ms < best) {
                        best = ms;
                        bestv = v;
                    }
                }
                variant[a] = bestv;
            }
            // Second pass, in context: a sweep timed alone leaves the caches
            // unlike a step does (the Picard layer parked in LDS, for one, wins
            // alone and loses between the other sweeps), so each axis's
            // variants are timed again over whole steps (2 Picard iterations:
            // both store forms of the X sweep), the other axes at their current
            // choice.  A step reads cur only; next is restored afterwards (its
            // NODE_OUT values may be rewritten by whole-run stores).
            {
                T* save = nullptr;
                CMCX_HIP(hipMalloc((void**)&save, sizeof(T) * 4 * NT));
                Layer4<T> sv;
                for (int f = 0; f < 4; f++) sv.f[f] = save + (int64_t)f * NT;
                copy_layer(sv, next);
                const bool short_prep = layers_ok;
                try {
                    for (int a : {AX_X, AX_Y, AX_Z}) {
                        if (plain_axis(a)) continue;
                        float best = 1e30f;
                        int bestv = variant[a];
                        for (int v = 0; v < (a == AX_Z ? kZVariants : kFusedVariants); v++) {
                            if (a != AX_Z && !variant_fits(a, v)) continue;
                            ensure_work(a, a == AX_Z ? 0 : v);
                            variant[a] = v;
                            enqueue_step(dt, 2, 1, short_prep);  // warm-up
                            CMCX_HIP(hipEventRecord(e0, st));
                            for (int r = 0; r < reps; r++) enqueue_step(dt, 2, 1, short_prep);
                            CMCX_HIP(hipEventRecord(e1, st));
                            CMCX_HIP(hipEventSynchronize(e1));
                            float ms = 0.f;
                            CMCX_HIP(hipEventElapsedTime(&ms, e0, e1));
                            if (ms < best) {
                                best = ms;
                                bestv = v;
                            }
                        }
                        variant[a] = bestv;
                        ensure_work(a, a == AX_Z ? 0 : bestv);
                    }
                } catch (...) {
                    copy_layer(next, sv);
                    hipStreamSynchronize(st);
                    hipFree(save);
                    throw;
                }
                copy_layer(next, sv);
                CMCX_HIP(hipStreamSynchronize(st));
                CMCX_HIP(hipFree(save));
            }
            CMCX_HIP(hipEventDestroy(e0));
            CMCX_HIP(hipEventDestroy(e1));
            tune_cache()[tune_key()] = {variant[0], variant[1], variant[2]};
            clear_graphs();
            layers_ok = false;  // the timed sweeps wrote half and tA
        }
        if (choice)
            for (int a = 0; a < 3; a++) choice[a] = variant[a];
        return CMCX_OK;
    }

    // Longest line of `axis` over the grids fits strided variant v.
    bool variant_fits(int axis, int v) const {
        for (const Grid& Gr : grids)
            if (!fused_variant_ok<T>(axis == AX_X ? Gr.g.nx : Gr.g.ny, v)) return false;
        return true;
    }

    int set_tile_variant(int axis, int v) override {
        if (axis < 0 || axis > 2 || v < 0 || v >= (axis == AX_Z ? kZVariants : kFusedVariants))
            return CMCX_ERR_ARG;
        if (axis != AX_Z && !variant_fits(axis, v)) {
            set_error("tile variant does not fit the line lengths of these grids");
            return CMCX_ERR_UNSUPPORTED;
        }
        variant[axis] = v;
        if (built) prepare_path_buffers();
        clear_graphs();
        return CMCX_OK;
    }

    int set_graph(int enable) override {
        graph_mode = enable != 0;
        if (!graph_mode) clear_graphs();
        return CMCX_OK;
    }

    int batch_size() const override { return (int)grids.size(); }

    int select(int index) override {
        if (index < 0 || index >= (int)grids.size()) return CMCX_ERR_ARG;
        sel = index;
        return CMCX_OK;
    }

    std::string tune_key() const {
        std::string k;
        for (const Grid& Gr : grids)
            k += std::to_string(Gr.g.nx) + "x" + std::to_string(Gr.g.ny) + "x" + std::to_string(Gr.g.nz) + ";";
        return k + "/" + std::to_string(sizeof(T));
    }

    int step_bytes(int G, int L, double* bytes) override {
        if (!bytes) return CMCX_ERR_ARG;
        if (!built) return CMCX_ERR_STATE;
        double S = 0;
        for (const Grid& Gr : grids) S += (double)Gr.segs[0].sum_size + Gr.segs[1].sum_size + Gr.segs[2].sum_size;
        *bytes = (double)G * L * (16.0 * sizeof(T) + 1.0) * S;
        return CMCX_OK;
    }

    int set_params(const cmcx_params* p) override {
        if (!p) return CMCX_ERR_ARG;
        G().prm = *p;
        invalidate_recording("set_params");
        upload_desc();
        clear_graphs();
        return CMCX_OK;
    }

    // ---- reverse mode ----------------------------------------------------------
    Layer4<T> lay(T* base, int64_t l) const {
        Layer4<T> r;
        for (int f = 0; f < 4; f++) r.f[f] = base + (l * 4 + f) * NT;
        return r;
    }

    void copy_layer(Layer4<T> dst, Layer4<T> src) { launch_copy4<T>(dst, src, NT, st); }

    int adjoint_forward(double dt, int steps, int G, int every) override {
        if (!built) return CMCX_ERR_STATE;
        if (steps < 1 || G < 1 || every < 1) return CMCX_ERR_ARG;
        if (grids.size() != 1 || merge_order != CMCX_MERGE_CPU) {
            set_error("adjoint: one grid with the CPU merge order");
            return CMCX_ERR_UNSUPPORTED;
        }
        if (!adjoint_supported<T>(grids[0].g) || grids[0].free_shared) {
            set_error("adjoint: axes of at most 256 nodes and no BC_FREE node shared by two segments");
            return CMCX_ERR_UNSUPPORTED;
        }
        const int nck = (steps + every - 1) / every;
        if (nck > adj.nck) {
            hipFree(adj.ck);
            adj.ck = nullptr;
            CMCX_HIP(hipMalloc((void**)&adj.ck, sizeof(T) * 4 * NT * nck));
            adj.nck = nck;
        }
        const T dtt = (T)dt;
        adj.dt = dtt;
        adj.G = G;
        adj.steps = steps;
        adj.every = every;
        ensure_tape(true);
        // the last intervals, one per tape buffer, are recorded on the way
        // (their layers go to the tape instead of the step buffers), so the
        // backward pass recomputes only the intervals before them, and starts
        // with two intervals to reverse while the third-last is recomputed
        const int nrec = std::min(nck, adj.nbuf);
        const int first0 = (nck - nrec) * every;
        const int64_t per = 1 + 6 * (int64_t)G;
        adj.holds[0] = adj.holds[1] = -1;
        for (int i = 0; i < steps; i++) {
            if (i % every == 0) copy_layer(lay(adj.ck, i / every), cur);
            if (i >= first0) {
                const int sgm = i / every;
                record_step(tape_buf(sgm % adj.nbuf) + (i - sgm * every) * per * 4 * NT);
            } else {
                step_keeping_out_fill(true, dtt, G, 1);
                std::swap(cur, next);
            }
        }
        for (int sgm = nck - nrec; sgm < nck; sgm++) adj.holds[sgm % adj.nbuf] = sgm;
        adj.recorded = true;
        adj.generation++;
        adj.invalid_why.clear();
        layers_ok = false;  // record_step leaves next and the Picard layers behind
        return CMCX_OK;
    }

    // Tape buffer b (of adj.nbuf), each holding one checkpoint interval.
    T* tape_buf(int b) const {
        const int64_t per = 1 + 6 * (int64_t)adj.G;
        return adj.tape + (int64_t)b * per * std::min(adj.every, adj.steps) * 4 * NT;
    }

    // Tape of the checkpoint intervals: two buffers when the rollout has more
    // than one interval (the backward pass recomputes one interval while it
    // reverses the next), else one.  The Picard-layer slots hold the node
    // values on non-IN nodes (the forward's temp holds the boundary-filled state
    // there, and the sweeps write NODE_IN nodes only).
    void ensure_tape(bool init_slots) {
        const int64_t per = 1 + 6 * (int64_t)adj.G;
        const int64_t nsteps_tape = std::min(adj.every, adj.steps);
        const int nbuf = adj.steps > adj.every ? 2 : 1;
        const int64_t need = per * nsteps_tape * nbuf;
        if (need > adj.tape_layers || adj.tape_G != adj.G) {
            hipFree(adj.tape);
            adj.tape = nullptr;
            CMCX_HIP(hipMalloc((void**)&adj.tape, sizeof(T) * 4 * NT * need));
            adj.tape_layers = need;
            adj.tape_G = adj.G;
            adj.holds[0] = adj.holds[1] = -1;
            init_slots = true;
        }
        adj.nbuf = nbuf;
        if (init_slots) {  // node values may have changed since the last recording
            Layer4<T> gl;
            for (int f = 0; f < 4; f++) gl.f[f] = d_g[f];
            for (int64_t i = 0; i < adj.tape_layers / per; i++)
                for (int k = 1; k <= 3 * adj.G; k += 2) {
                    Layer4<T> a = lay(adj.tape + i * per * 4 * NT, k);
                    Layer4<T> b = k + 1 <= 3 * adj.G ? lay(adj.tape + i * per * 4 * NT, k + 1) : a;
                    launch_prep_step_list<T>(n_nonin, d_nonin, d_type, gl, a, a, b, st);
                }
        }
    }

    // UpdateBoundaries + one step from `cur`, keeping every layer the reverse
    // step reads in tape slot base sb; leaves the new state in cur.
    void record_step(T* sb) {
        const int G = adj.G;
        const T dt = adj.dt;
        update_boundaries();
        Layer4<T> S = lay(sb, 0);
        copy_layer(S, cur);
        Layer4<T> tin = S;
        const int axes[3] = {AX_Z, AX_Y, AX_X};
        for (int it = 0; it < G; it++)
            for (int q = 0; q < 3; q++) {
                const int k = 3 * it + q;
                Layer4<T> in = q == 0 ? S : lay(sb, 1 + 3 * G + k - 1);
                Layer4<T> out = lay(sb, 1 + 3 * G + k), tout = lay(sb, k + 1);
                // the step's Z and X sweeps share one output buffer, which holds
                // the state off the segments at the start of the step: the last
                // Z sweep starts from the state and the last X sweep from the Z
                // sweep's output, whose values they keep on the nodes they do
                // not cover (the X output becomes the new state)
                if (q == 0 && it == G - 1) copy_layer(out, S);
                if (q == 2 && it == G - 1) copy_layer(out, lay(sb, 1 + 3 * G + 3 * it));
                sweep(axes[q], in, tin, out, tout, q == 2 ? 2 : 1, 0, 4, dt);
                tin = tout;
            }
        copy_layer(cur, lay(sb, 1 + 3 * G + 3 * G - 1));
    }

    void reverse_step(T* sb) {
        const int G = adj.G;
        const Geom& g = grids[0].g;
        Layer4<T> XB = lay(adj.work, 0), TB = lay(adj.work, 1), CB = lay(adj.work, 2),
                  LB = lay(adj.work, 3), GB = lay(adj.work, 4);
        const int axes[3] = {AX_Z, AX_Y, AX_X};
        for (int it = G - 1; it >= 0; it--)
            for (int q = 2; q >= 0; q--) {
                const int k = 3 * it + q, a = axes[q];
                Layer4<T> tin = lay(sb, k == 0 ? 0 : k), out = lay(sb, 1 + 3 * G + k);
                const T alpha = q == 2 ? (T)0.25 : (T)0.5, beta = q == 2 ? (T)0.75 : (T)0.5;
                const SweepParams<T> sp = sweep_params(0, a, adj.dt);
                // adjoint of the sweep output: the step's seed (XB) for its last
                // X sweep, zero for the other X sweeps (their output feeds only
                // the Picard layer), (3/dt) lam of the sweep reversed before
                const Layer4<T> xs = q == 2 ? (it == G - 1 ? XB : Layer4<T>{}) : LB;
                const T xscale = q == 2 ? (T)1 : sp.three_dt;
                launch_adj_sweep<T>(g, a, sp, d_type, d_bcv, d_bct, tin, out, xs, xscale, XB, TB, CB,
                                    LB, GB, alpha, beta, q == 0 ? 1 : 0,
                                    (q == 0 && it == G - 1) ? 1 : 0, (q == 0 && it == 0) ? 1 : 0,
                                    adj.tiles[a], adj.ntiles[a], adj.ltiles[a], adj.pp, adj.nmask,
                                    st);
            }
    }

    int adjoint_backward(const void* const* seed, void* const* gbar, void* const* sbar,
                         double* pbar, bool device) override {
        if (!adj.recorded) {
            set_error(adj.invalid_why.empty()
                          ? std::string("adjoint_backward before adjoint_forward")
                          : "the recorded rollout was invalidated by " + adj.invalid_why +
                                " (record it again with adjoint_forward)");
            return CMCX_ERR_STATE;
        }
        const Geom& g = grids[0].g;
        const int64_t per = 1 + 6 * (int64_t)adj.G;
        ensure_tape(false);
        if (!adj.work) {
            CMCX_HIP(hipMalloc((void**)&adj.work, sizeof(T) * 4 * NT * 6));
            CMCX_HIP(hipMalloc((void**)&adj.pp, sizeof(double) * 4 * adj_point_blocks()));
            CMCX_HIP(hipMalloc((void**)&adj.nmask, sizeof(uint16_t) * g.N));
            for (int a = 0; a < 3; a++)
                build_adj_tiles(h_type.data(), g, a, adj_lpg<T>(g), &adj.tiles[a], &adj.ntiles[a],
                                &adj.ltiles[a]);
        }
        Layer4<T> XB = lay(adj.work, 0), GB = lay(adj.work, 4), FIN = lay(adj.work, 5);
        copy_layer(FIN, cur);
        // the recomputed intervals start from checkpoints taken before their
        // boundary fill: full fills there; the state and its flags come back at the end
        const bool synced0 = bv_synced, layers0 = layers_ok;
        CMCX_HIP(hipMemsetAsync(adj.work, 0, sizeof(T) * 4 * NT * 5, st));
        CMCX_HIP(hipMemsetAsync(adj.pp, 0, sizeof(double) * 4 * adj_point_blocks(), st));
        // the BC kinds may have changed since the last backward (set_boundary)
        launch_adj_mask(g, d_type, d_bcv, d_bct, adj.nmask, st);
        if (seed)
            for (int f = 0; f < 4; f++)
                if (seed[f])
                    CMCX_HIP(hipMemcpyAsync(XB.f[f], seed[f], sizeof(T) * g.N,
                                            device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
        // Intervals are reversed last to first.  Interval i - 1 is recomputed
        // (from its checkpoint, into the other tape buffer, on a stream of its
        // own) while interval i is reversed on the solver's stream; events
        // order a reverse after its interval's recompute, and a recompute
        // after the reverse that last used its buffer.
        const int nseg = (adj.steps + adj.every - 1) / adj.every;
        const int nb = adj.nbuf;
        if (!adj.fs) CMCX_HIP(hipStreamCreateWithFlags(&adj.fs, hipStreamNonBlocking));
        std::vector<hipEvent_t> recd(nseg, nullptr), revd(nseg, nullptr);
        hipEvent_t start = nullptr;
        struct EvFree {
            std::vector<hipEvent_t>* a;
            std::vector<hipEvent_t>* b;
            hipEvent_t* c;
            ~EvFree() {
                for (auto* v : {a, b})
                    for (hipEvent_t e : *v)
                        if (e) hipEventDestroy(e);
                if (*c) hipEventDestroy(*c);
            }
        } evfree{&recd, &revd, &start};
        for (int i = 0; i < nseg; i++) {
            CMCX_HIP(hipEventCreateWithFlags(&recd[i], hipEventDisableTiming));
            CMCX_HIP(hipEventCreateWithFlags(&revd[i], hipEventDisableTiming));
        }
        CMCX_HIP(hipEventCreateWithFlags(&start, hipEventDisableTiming));
        CMCX_HIP(hipEventRecord(start, st));
        CMCX_HIP(hipStreamWaitEvent(adj.fs, start, 0));
        hipStream_t main_st = st;
        std::vector<bool> recomputed(nseg, false);
        auto recompute = [&](int sgm) {
            const int k = std::min(adj.every, adj.steps - sgm * adj.every);
            const int b = sgm % nb;
            if (sgm + 2 < nseg) CMCX_HIP(hipStreamWaitEvent(adj.fs, revd[sgm + 2], 0));
            st = adj.fs;
            try {
                copy_layer(cur, lay(adj.ck, sgm));
                bv_synced = false;
                for (int i = 0; i < k; i++) record_step(tape_buf(b) + i * per * 4 * NT);
            } catch (...) {
                st = main_st;
                throw;
            }
            st = main_st;
            CMCX_HIP(hipEventRecord(recd[sgm], adj.fs));
            adj.holds[b] = sgm;
            recomputed[sgm] = true;
        };
        if (adj.holds[(nseg - 1) % nb] != nseg - 1) recompute(nseg - 1);
        for (int sgm = nseg - 1; sgm >= 0; sgm--) {
            if (sgm > 0 && adj.holds[(sgm - 1) % nb] != sgm - 1) {
                if (nb == 1) {
                    // one buffer: the recompute waits for this interval's reverse
                    // (single-interval rollouts never get here)
                } else {
                    recompute(sgm - 1);
                }
            }
            const int k = std::min(adj.every, adj.steps - sgm * adj.every);
            if (recomputed[sgm]) CMCX_HIP(hipStreamWaitEvent(st, recd[sgm], 0));
            for (int i = k - 1; i >= 0; i--) reverse_step(tape_buf(sgm % nb) + i * per * 4 * NT);
            CMCX_HIP(hipEventRecord(revd[sgm], st));
        }
        if (gbar)
            for (int f = 0; f < 4; f++)
                if (gbar[f])
                    CMCX_HIP(hipMemcpyAsync(gbar[f], GB.f[f], sizeof(T) * g.N,
                                            device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
        if (sbar)
            for (int f = 0; f < 4; f++)
                if (sbar[f])
                    CMCX_HIP(hipMemcpyAsync(sbar[f], XB.f[f], sizeof(T) * g.N,
                                            device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
        std::vector<double> pp(4 * adj_point_blocks());
        CMCX_HIP(hipMemcpyAsync(pp.data(), adj.pp, sizeof(double) * pp.size(), hipMemcpyDeviceToHost, st));
        copy_layer(cur, FIN);
        bv_synced = synced0;
        layers_ok = layers0;
        CMCX_HIP(hipStreamSynchronize(st));
        if (pbar) {
            for (int q = 0; q < 4; q++) pbar[q] = 0;
            for (int b = 0; b < adj_point_blocks(); b++)
                for (int q = 0; q < 4; q++) pbar[q] += pp[4 * b + q];
        }
        return CMCX_OK;
    }

    // The recompute and reverse of a recorded rollout read the solver's
    // parameters, node values, BC kinds and path: changing one of them leaves a
    // recording whose gradient would belong to another rollout.
    void invalidate_recording(const char* why) {
        if (!adj.recorded) return;
        adj.recorded = false;
        adj.generation++;
        adj.invalid_why = why;
    }

    int adjoint_generation(uint64_t* gen) override {
        if (!gen) return CMCX_ERR_ARG;
        *gen = adj.recorded ? adj.generation : 0;
        return CMCX_OK;
    }

    int adjoint_release() override {
        hipStreamSynchronize(st);
        hipFree(adj.ck);
        hipFree(adj.tape);
        adj.ck = nullptr;
        adj.tape = nullptr;
        adj.nck = 0;
        adj.tape_layers = 0;
        adj.holds[0] = adj.holds[1] = -1;
        invalidate_recording("adjoint_release");
        return CMCX_OK;
    }

    int synchronize() override {
        CMCX_HIP(hipStreamSynchronize(st));
        return CMCX_OK;
    }
};

}  // namespace cmcx

using namespace cmcx;

struct cmcx_solver {
    int dtype;
    std::unique_ptr<ISolver> impl;
};

#define CMCX_GUARD(body)                                                                       \
    try {                                                                                      \
        body                                                                                   \
    } catch (const cmcx::HipError& e) {                                                        \
        cmcx::set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where); \
        return CMCX_ERR_HIP;                                                                   \
    } catch (const std::exception& e) {                                                        \
        cmcx::set_error(e.what());                                                             \
        return CMCX_ERR_ARG;                                                                   \
    } catch (const std::string& e) {                                                           \
        cmcx::set_error(e);                                                                    \
        return CMCX_ERR_ARG;                                                                   \
    }

#define SOLVER_CALL(s, expr)                  \
    do {                                      \
        if (!(s) || !(s)->impl) return CMCX_ERR_ARG; \
        CMCX_GUARD(return (s)->impl->expr;)   \
    } while (0)

extern "C" {

const char* cmcx_version(void) { return "cmcx 0.1 (gfx950)"; }
const char* cmcx_last_error(void) { return cmcx::g_last_error.c_str(); }

void cmcx_params_normalized(double Re, double Pr, double lambda, cmcx_params* out) {
    if (!out) return;
    out->v_T = 1.0;
    out->v_vis = 1.0 / Re;
    out->t_vis = 1.0 / (Re * Pr);
    out->t_phi = (lambda - 1) / (lambda * Re);
}

void cmcx_params_dimensional(double vis, double rho, double R, double k, double cv,
                             cmcx_params* out) {
    if (!out) return;
    out->v_T = R;
    out->v_vis = vis / rho;
    out->t_vis = k / (rho * cv);
    out->t_phi = vis / (rho * cv);
}

static int create_solver(const cmcx_domain* doms, const cmcx_params* params, int count, int dtype,
                         void* stream, cmcx_solver** out) {
    if (!doms || !params || !out || count < 1) return CMCX_ERR_ARG;
    *out = nullptr;
    if (dtype != CMCX_F32 && dtype != CMCX_F64) return CMCX_ERR_ARG;
    for (int i = 0; i < count; i++) {
        const cmcx_domain* dom = &doms[i];
        for (int a = 0; a < 3; a++)
            if (dom->dims[a] < 3) {
                set_error("every dimension must be >= 3");
                return CMCX_ERR_ARG;
            }
        if (!dom->type || !dom->bc_vel || !dom->bc_temp || !dom->v[0] || !dom->v[1] ||
            !dom->v[2] || !dom->T)
            return CMCX_ERR_ARG;
        int64_t N = (int64_t)dom->dims[0] * dom->dims[1] * dom->dims[2];
        if (N > ((int64_t)1 << 31) - 1) {
            set_error("grid too large (node count must fit in int32)");
            return CMCX_ERR_TOO_LARGE;
        }
        for (int64_t k = 0; k < N; k++)
            if (dom->type[k] > 3 || dom->bc_vel[k] > 1 || dom->bc_temp[k] > 1) {
                set_error("node type or BC kind out of range at node " + std::to_string(k) +
                          " of grid " + std::to_string(i));
                return CMCX_ERR_ARG;
            }
    }
    CMCX_GUARD({
        auto* s = new cmcx_solver;
        s->dtype = dtype;
        hipStream_t st = (hipStream_t)stream;
        try {
            if (dtype == CMCX_F32) s->impl.reset(new Solver<float>(doms, params, count, st));
            else s->impl.reset(new Solver<double>(doms, params, count, st));
        } catch (const std::string& e) {
            delete s;
            set_error(e);
            return CMCX_ERR_TOO_LARGE;
        } catch (...) {
            delete s;
            throw;
        }
        int rc = dtype == CMCX_F32 ? static_cast<Solver<float>*>(s->impl.get())->check_mask()
                                   : static_cast<Solver<double>*>(s->impl.get())->check_mask();
        if (rc != CMCX_OK) {
            delete s;
            return rc;
        }
        *out = s;
        return CMCX_OK;
    })
}

int cmcx_create(const cmcx_domain* dom, const cmcx_params* params, int dtype, void* stream,
                cmcx_solver** out) {
    return create_solver(dom, params, 1, dtype, stream, out);
}

int cmcx_create_batch(const cmcx_domain* doms, const cmcx_params* params, int count, int dtype,
                      void* stream, cmcx_solver** out) {
    return create_solver(doms, params, count, dtype, stream, out);
}

int cmcx_destroy(cmcx_solver* s) {
    if (!s) return CMCX_ERR_ARG;
    CMCX_GUARD({
        delete s;
        return CMCX_OK;
    })
}

int cmcx_set_options(cmcx_solver* s, int path, int merge_order) { SOLVER_CALL(s, set_options(path, merge_order)); }
int cmcx_set_nblockz(cmcx_solver* s, int nb) { SOLVER_CALL(s, set_nblockz(nb)); }
int cmcx_build_segments(cmcx_solver* s) { SOLVER_CALL(s, build_segments()); }
int cmcx_num_segments(cmcx_solver* s, int axis, int64_t* n) { SOLVER_CALL(s, num_segments(axis, n)); }
int cmcx_get_segments(cmcx_solver* s, int axis, int32_t* out, int64_t cap) { SOLVER_CALL(s, get_segments(axis, out, cap)); }
int cmcx_get_segment_nodes(cmcx_solver* s, int axis, uint8_t* flags, double* vals, int64_t cap) {
    SOLVER_CALL(s, get_segment_nodes(axis, flags, vals, cap));
}
int cmcx_update_boundaries(cmcx_solver* s) { SOLVER_CALL(s, update_boundaries()); }
int cmcx_step(cmcx_solver* s, double dt, int G, int L, int compute_err, double* err) {
    SOLVER_CALL(s, step(dt, G, L, compute_err, err));
}
int cmcx_run(cmcx_solver* s, double dt, int steps, int G, int L, double* err) {
    SOLVER_CALL(s, run(dt, steps, G, L, err));
}
int cmcx_div_error(cmcx_solver* s, double* err) { SOLVER_CALL(s, div_error(err)); }
int cmcx_get_fields(cmcx_solver* s, void* u, void* v, void* w, void* p, void* T) {
    SOLVER_CALL(s, get_fields(u, v, w, p, T, false));
}
int cmcx_get_fields_device(cmcx_solver* s, void* u, void* v, void* w, void* p, void* T) {
    SOLVER_CALL(s, get_fields(u, v, w, p, T, true));
}
int cmcx_set_fields(cmcx_solver* s, const void* u, const void* v, const void* w, const void* T) {
    SOLVER_CALL(s, set_fields(u, v, w, T));
}
int cmcx_state_device_ptrs(cmcx_solver* s, void** p4) { SOLVER_CALL(s, state_ptrs(p4)); }
int cmcx_get_layer(cmcx_solver* s, void* vel, double* T, int ox, int oy, int oz) {
    SOLVER_CALL(s, get_layer(vel, T, ox, oy, oz));
}
int cmcx_set_boundary(cmcx_solver* s, const uint8_t* bv, const uint8_t* bt, const double* vx,
                      const double* vy, const double* vz, const double* T) {
    SOLVER_CALL(s, set_boundary(bv, bt, vx, vy, vz, T));
}
int cmcx_profile(cmcx_solver* s, int enable) { SOLVER_CALL(s, profile(enable)); }
int cmcx_profile_read(cmcx_solver* s, double* ms, int64_t* launches, int reset) {
    SOLVER_CALL(s, profile_read(ms, launches, reset));
}
int cmcx_step_bytes(cmcx_solver* s, int G, int L, double* bytes) { SOLVER_CALL(s, step_bytes(G, L, bytes)); }
int cmcx_autotune(cmcx_solver* s, int reps, int* choice) { SOLVER_CALL(s, autotune(reps, choice)); }
int cmcx_set_tile_variant(cmcx_solver* s, int axis, int variant) {
    SOLVER_CALL(s, set_tile_variant(axis, variant));
}
int cmcx_set_graph(cmcx_solver* s, int enable) { SOLVER_CALL(s, set_graph(enable)); }
int cmcx_batch_size(cmcx_solver* s, int* count) {
    if (!s || !s->impl || !count) return CMCX_ERR_ARG;
    *count = s->impl->batch_size();
    return CMCX_OK;
}
int cmcx_select(cmcx_solver* s, int index) { SOLVER_CALL(s, select(index)); }
int cmcx_grid_errors(cmcx_solver* s, double* errs) { SOLVER_CALL(s, grid_errors(errs)); }
int cmcx_set_params(cmcx_solver* s, const cmcx_params* p) { SOLVER_CALL(s, set_params(p)); }
int cmcx_adjoint_forward(cmcx_solver* s, double dt, int steps, int num_global, int checkpoint_every) {
    SOLVER_CALL(s, adjoint_forward(dt, steps, num_global, checkpoint_every));
}
int cmcx_adjoint_backward(cmcx_solver* s, const void* const* seed, void* const* gbar,
                          void* const* sbar, double* pbar) {
    SOLVER_CALL(s, adjoint_backward(seed, gbar, sbar, pbar, false));
}
int cmcx_adjoint_backward_device(cmcx_solver* s, const void* const* seed, void* const* gbar,
                                 void* const* sbar, double* pbar) {
    SOLVER_CALL(s, adjoint_backward(seed, gbar, sbar, pbar, true));
}
int cmcx_adjoint_release(cmcx_solver* s) { SOLVER_CALL(s, adjoint_release()); }
int cmcx_adjoint_generation(cmcx_solver* s, uint64_t* gen) { SOLVER_CALL(s, adjoint_generation(gen)); }
int cmcx_synchronize(cmcx_solver* s) { SOLVER_CALL(s, synchronize()); }

}  // extern "C"
