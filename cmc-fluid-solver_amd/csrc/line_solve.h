// Device building blocks of the line-wide tridiagonal solves (forward sweeps
// and their adjoint): row roles, the chunked partition elimination and the
// solve of the reduced systems.
//
// Chunked partition method (the Thomas/PCR hybrid of Kim, Wu, Chang and Hwu,
// "A scalable tridiagonal solver for GPUs", 2011): the rows of a line are cut
// into chunks of M consecutive rows held by one lane.  A downward then upward
// elimination inside the chunk expresses every inner row through the chunk's
// first and last unknowns, x_j = E_j + F_j x_first + G_j x_last; the first and
// last rows of all chunks form a tridiagonal "reduced" system of 2 rows per
// chunk.  One cyclic-reduction step eliminates the first unknowns (one row per
// chunk remains), parallel cyclic reduction with one row per lane and
// cross-lane shuffles solves for the last unknowns, the first ones follow from
// their rows, and each lane then reconstructs its inner rows.  The reduced rows
// reach the solving lanes through LDS when a line's chunks span wavefronts,
// and stay in registers when they are the lanes of one wavefront.  Two
// matrices share the work: matrix 0 (velocity) serves right-hand sides 0..2,
// matrix 1 (T) right-hand side 3.
#pragma once

#include "cmcx_internal.h"

#include <type_traits>

namespace cmcx {
namespace lsolve {

enum { RL_NONE = 0, RL_IN = 1, RL_START = 2, RL_END = 3, RL_SHARED = 4 };
constexpr int kRedFields = 10;  // A_v B_v C_v A_t B_t C_t D_u D_v D_w D_t
constexpr uint8_t kOutside = 255;

__device__ __forceinline__ int row_role(uint8_t t, uint8_t tl, uint8_t tr) {
    if (t == NODE_IN) return RL_IN;
    bool L = tl == NODE_IN, R = tr == NODE_IN;
    if (L && R) return RL_SHARED;
    if (R) return RL_START;
    if (L) return RL_END;
    return RL_NONE;
}

// Boundary or identity row of one matrix (ApplyBC0/ApplyBC1); uses_val: the
// RHS is the node value.
template <class T>
__device__ __forceinline__ void bc_row(int role, bool free_bc, T& a, T& b, T& c, bool& uses_val) {
    uses_val = false;
    if (role == RL_START && free_bc) { a = 0; b = 2; c = -1; return; }
    if (role == RL_END && free_bc) { a = -1; b = 2; c = 0; return; }
    a = 0; b = 1; c = 0;
    uses_val = role != RL_NONE;
}

inline __device__ unsigned xcd_remap(unsigned b, unsigned nb) {
    const unsigned X = 8;
    unsigned q = nb / X, rem = nb % X;
    unsigned xcd = b % X, local = b / X;
    return xcd < rem ? xcd * (q + 1) + local : rem * (q + 1) + (xcd - rem) * q + local;
}

// Reciprocal used by the eliminations: FP32 takes the hardware reciprocal
// (v_rcp_f32, 1 ulp) instead of the ~10-instruction correctly rounded divide;
// FP64 keeps the exact division.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp(double x) { return 1.0 / x; }

template <class T>
struct Row {
    T av, bv, cv, at, bt, ct;
    T d[4];
};

// Chunk state of one lane: inner rows 1..M-2 as x_j = E_j + F_j x_0 + G_j x_{M-1}
// (F, G per matrix, E per field).
template <class T, int M>
struct Chunk {
    T Fv[M], Gv[M], Ft[M], Gt[M];
    T E[4][M];
};

// Downward elimination, row j in 1..M-2: x_j = p_j + q_j x_0 - r_j x_{j+1},
// (p, q, r) kept in (E, F, G).
template <class T, int M>
__device__ __forceinline__ void chunk_down(Chunk<T, M>& ch, int j, const Row<T>& r) {
    if (j == 1) {
        T iv = rcp(r.bv), it = rcp(r.bt);
        ch.Fv[j] = -r.av * iv; ch.Gv[j] = r.cv * iv;
        ch.Ft[j] = -r.at * it; ch.Gt[j] = r.ct * it;
#pragma unroll
        for (int v = 0; v < 3; v++) ch.E[v][j] = r.d[v] * iv;
        ch.E[3][j] = r.d[3] * it;
    } else {
        T iv = rcp(r.bv - r.av * ch.Gv[j - 1]);
        T it = rcp(r.bt - r.at * ch.Gt[j - 1]);
#pragma unroll
        for (int v = 0; v < 3; v++) ch.E[v][j] = (r.d[v] - r.av * ch.E[v][j - 1]) * iv;
        ch.E[3][j] = (r.d[3] - r.at * ch.E[3][j - 1]) * it;
        ch.Fv[j] = -r.av * ch.Fv[j - 1] * iv;
        ch.Ft[j] = -r.at * ch.Ft[j - 1] * it;
        ch.Gv[j] = r.cv * iv;
        ch.Gt[j] = r.ct * it;
    }
}

// Upward elimination: (p, q, r) -> (E, F, G).
template <class T, int M>
__device__ __forceinline__ void chunk_up(Chunk<T, M>& ch) {
    ch.Gv[M - 2] = -ch.Gv[M - 2];
    ch.Gt[M - 2] = -ch.Gt[M - 2];
#pragma unroll
    for (int j = M - 3; j >= 1; j--) {
        T rv = ch.Gv[j], rt = ch.Gt[j];
#pragma unroll
        for (int v = 0; v < 3; v++) ch.E[v][j] = ch.E[v][j] - rv * ch.E[v][j + 1];
        ch.E[3][j] = ch.E[3][j] - rt * ch.E[3][j + 1];
        ch.Fv[j] = ch.Fv[j] - rv * ch.Fv[j + 1];
        ch.Ft[j] = ch.Ft[j] - rt * ch.Ft[j + 1];
        ch.Gv[j] = -rv * ch.Gv[j + 1];
        ch.Gt[j] = -rt * ch.Gt[j + 1];
    }
}

// Reduced rows of a chunk: f couples (x_{-1}, x_0, x_{M-1}), l couples
// (x_0, x_{M-1}, x_M).  Fields A_v B_v C_v A_t B_t C_t D_u D_v D_w D_t.
template <class T, int M>
__device__ __forceinline__ void chunk_reduced(const Chunk<T, M>& ch, const Row<T>& r0,
                                              const Row<T>& rl, T* f, T* l) {
    if (M >= 3) {
        f[0] = r0.av; f[1] = r0.bv + r0.cv * ch.Fv[1]; f[2] = r0.cv * ch.Gv[1];
        f[3] = r0.at; f[4] = r0.bt + r0.ct * ch.Ft[1]; f[5] = r0.ct * ch.Gt[1];
#pragma unroll
        for (int v = 0; v < 3; v++) f[6 + v] = r0.d[v] - r0.cv * ch.E[v][1];
        f[9] = r0.d[3] - r0.ct * ch.E[3][1];
        l[0] = rl.av * ch.Fv[M - 2]; l[1] = rl.bv + rl.av * ch.Gv[M - 2]; l[2] = rl.cv;
        l[3] = rl.at * ch.Ft[M - 2]; l[4] = rl.bt + rl.at * ch.Gt[M - 2]; l[5] = rl.ct;
#pragma unroll
        for (int v = 0; v < 3; v++) l[6 + v] = rl.d[v] - rl.av * ch.E[v][M - 2];
        l[9] = rl.d[3] - rl.at * ch.E[3][M - 2];
    } else {
        f[0] = r0.av; f[1] = r0.bv; f[2] = r0.cv; f[3] = r0.at; f[4] = r0.bt; f[5] = r0.ct;
        l[0] = rl.av; l[1] = rl.bv; l[2] = rl.cv; l[3] = rl.at; l[4] = rl.bt; l[5] = rl.ct;
#pragma unroll
        for (int v = 0; v < 4; v++) { f[6 + v] = r0.d[v]; l[6 + v] = rl.d[v]; }
    }
}

// Values of the neighbouring lanes of the wavefront (lane - 1, lane + 1) by
// DPP wave shifts, which move data in the VALU instead of through the LDS
// crossbar like __shfl; lane 0 (lane 63) receives 0.  Called with every lane
// of the wavefront active, as __shfl is here.
__device__ __forceinline__ int lane_prev(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int lane_next(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ float lane_prev(float x) { return __int_as_float(lane_prev(__float_as_int(x))); }
__device__ __forceinline__ float lane_next(float x) { return __int_as_float(lane_next(__float_as_int(x))); }
__device__ __forceinline__ double lane_prev(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)lane_prev((int)(unsigned)b), hi = (unsigned)lane_prev((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double lane_next(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)lane_next((int)(unsigned)b), hi = (unsigned)lane_next((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Parallel cyclic reduction over one reduced system per group of RP lanes:
// the lane holds row i = lane % RP (rows >= R are identity, and every access
// stays inside the lane's group).  On return D holds x.
template <class T>
__device__ __forceinline__ void pcr_solve(T* A, T* B, T* C, T* D, int i, int R, int RP) {
    // A[q], B[q], C[q]: matrix q (0 vel, 1 T); D[0..2] use matrix 0, D[3] matrix 1.
    // Rows are kept divided by their diagonal (B = 1), so a level exchanges A, C
    // and D only and takes one reciprocal per matrix.
    if (i >= R) {
        A[0] = A[1] = C[0] = C[1] = 0;
        B[0] = B[1] = 1;
        D[0] = D[1] = D[2] = D[3] = 0;
    }
    if (i == 0) A[0] = A[1] = 0;
    if (i == R - 1) C[0] = C[1] = 0;
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const T ib = rcp(B[q]);
        A[q] = A[q] * ib;
        C[q] = C[q] * ib;
#pragma unroll
        for (int v = (q == 0 ? 0 : 3); v < (q == 0 ? 3 : 4); v++) D[v] = D[v] * ib;
    }
    const int lane = threadIdx.x & 63;
    for (int st = 1; st < RP; st <<= 1) {
        const bool hm = i - st >= 0, hp = i + st < RP;
        const int lm = hm ? lane - st : lane, lp = hp ? lane + st : lane;
        T Am[2], Cm[2], Ap[2], Cp[2], Dm[4], Dp[4];
        if (st == 1) {
#pragma unroll
            for (int q = 0; q < 2; q++) {
                Am[q] = lane_prev(A[q]); Cm[q] = lane_prev(C[q]);
                Ap[q] = lane_next(A[q]); Cp[q] = lane_next(C[q]);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) { Dm[v] = lane_prev(D[v]); Dp[v] = lane_next(D[v]); }
        } else {
#pragma unroll
            for (int q = 0; q < 2; q++) {
                Am[q] = __shfl(A[q], lm, 64); Cm[q] = __shfl(C[q], lm, 64);
                Ap[q] = __shfl(A[q], lp, 64); Cp[q] = __shfl(C[q], lp, 64);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) { Dm[v] = __shfl(D[v], lm, 64); Dp[v] = __shfl(D[v], lp, 64); }
        }
        if (!hm) {
#pragma unroll
            for (int q = 0; q < 2; q++) { Am[q] = 0; Cm[q] = 0; }
#pragma unroll
            for (int v = 0; v < 4; v++) Dm[v] = 0;
        }
        if (!hp) {
#pragma unroll
            for (int q = 0; q < 2; q++) { Ap[q] = 0; Cp[q] = 0; }
#pragma unroll
            for (int v = 0; v < 4; v++) Dp[v] = 0;
        }
#pragma unroll
        for (int q = 0; q < 2; q++) {
            // row i - A row(i - st) - C row(i + st), divided again by its diagonal
            const T ib = rcp((T)1 - Cm[q] * A[q] - Ap[q] * C[q]);
#pragma unroll
            for (int v = (q == 0 ? 0 : 3); v < (q == 0 ? 3 : 4); v++)
                D[v] = (D[v] - Dm[v] * A[q] - Dp[v] * C[q]) * ib;
            const T a = -Am[q] * A[q], c = -Cp[q] * C[q];
            A[q] = a * ib;
            C[q] = c * ib;
        }
    }
}

// Identity reduced row (x = 0) in the kRedFields layout.
template <class T>
__device__ __forceinline__ void ident_row(T* r) {
#pragma unroll
    for (int k = 0; k < kRedFields; k++) r[k] = (k == 1 || k == 4) ? (T)1 : (T)0;
}

// One cyclic-reduction step on the reduced system (2 rows per chunk): chunk i's
// rows f_i (x_{l,i-1}, x_{f,i}, x_{l,i}) and l_i (x_{f,i}, x_{l,i}, x_{f,i+1}),
// with the next chunk's f_{i+1}, give after eliminating x_{f,i} and x_{f,i+1}
// row i of a tridiagonal system in the chunks' last unknowns:
//   A = -m a_f,  B = b_l - m c_f - m' a_f',  C = -m' c_f',  D = d_l - m d_f - m' d_f'
// with m = a_l / b_f and m' = c_l / b_f' (primes: chunk i+1, an identity row
// when there is none).  ib receives 1 / b_f per matrix for first_unknowns.
template <class T>
__device__ __forceinline__ void compact_rows(const T* f, const T* l, const T* fn, T* A, T* B,
                                             T* C, T* D, T* ib) {
#pragma unroll
    for (int q = 0; q < 2; q++) {
        ib[q] = rcp(f[3 * q + 1]);
        const T m = l[3 * q] * ib[q];
        const T mn = l[3 * q + 2] * rcp(fn[3 * q + 1]);
        A[q] = -m * f[3 * q];
        B[q] = l[3 * q + 1] - m * f[3 * q + 2] - mn * fn[3 * q];
        C[q] = -mn * fn[3 * q + 2];
#pragma unroll
        for (int v = (q == 0 ? 0 : 3); v < (q == 0 ? 3 : 4); v++)
            D[v] = l[6 + v] - m * f[6 + v] - mn * fn[6 + v];
    }
}

// x_{f,i} from row f_i once the last unknowns of chunks i-1 (xlp) and i (xl)
// are known.
template <class T>
__device__ __forceinline__ void first_unknowns(const T* f, const T* ib, const T* xlp, const T* xl,
                                               T* xf) {
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int q = v < 3 ? 0 : 1;
        xf[v] = (f[6 + v] - f[3 * q] * xlp[v] - f[3 * q + 2] * xl[v]) * ib[q];
    }
}

// Reduced system of one line held in registers by RP consecutive lanes (a power
// of two; lane i of the group holds chunk i's rows f, l; R chunks, identity rows
// beyond): compaction to the last unknowns, PCR across the lanes, then the first
// unknowns.  Every lane of the wavefront must call it.
template <class T>
__device__ __forceinline__ void solve_chunks_in_lanes(const T* f, const T* l, int i, int R, int RP,
                                                      T* xf, T* xl) {
    const int lane = threadIdx.x & 63;
    const bool hn = i + 1 < R;
    T fn[kRedFields];
#pragma unroll
    for (int k = 0; k < kRedFields; k++) {
        const T v = lane_next(f[k]);
        fn[k] = hn ? v : ((k == 1 || k == 4) ? (T)1 : (T)0);
    }
    T A[2], B[2], C[2], D[4], ib[2];
    compact_rows<T>(f, l, fn, A, B, C, D, ib);
    pcr_solve<T>(A, B, C, D, i, R, RP);
    T xlp[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const T t = lane_prev(D[v]);
        xlp[v] = i > 0 ? t : (T)0;
        xl[v] = D[v];
    }
    first_unknowns<T>(f, ib, xlp, xl, xf);
}

// Reduced systems staged in LDS as [line][field][row] with a line pitch of
// kRedFields*R + 1 elements (conflict-free for both the chunk writes, one lane
// per line, and the PCR reads, one lane per row).
struct RedLayout {
    int R, pitch;
    __device__ __forceinline__ int at(int line, int field, int row) const {
        return line * pitch + field * R + row;
    }
};

// Solves `nl` reduced systems held in LDS (row 2i = f_i, row 2i+1 = l_i of
// chunk i, L.R / 2 chunks per line) with the lanes of `nw` wavefronts (this one
// is wave `w`): compaction to one unknown per chunk, one lane per chunk with the
// lanes of a line inside one wavefront, PCR over L.R / 2 rows; writes x_f and
// x_l over the D fields of rows 2i and 2i+1.
template <class T>
__device__ __forceinline__ void solve_reduced_compact(T* red, RedLayout L, int nl, int w, int nw) {
    const int lane = threadIdx.x & 63;
    const int P = L.R / 2;
    int RP = 1;
    while (RP < P) RP <<= 1;
    const int lpp = 64 / RP;  // lines per pass of one wave
    const int i = lane % RP;
    for (int base = w * lpp; base < nl; base += nw * lpp) {
        const int line = base + lane / RP;
        const int Rl = line < nl ? P : 0;
        const bool act = i < Rl, actn = i + 1 < Rl;
        T f[kRedFields], l[kRedFields], fn[kRedFields];
#pragma unroll
        for (int k = 0; k < kRedFields; k++) {
            const T id = (k == 1 || k == 4) ? (T)1 : (T)0;
            f[k] = act ? red[L.at(line, k, 2 * i)] : id;
            l[k] = act ? red[L.at(line, k, 2 * i + 1)] : id;
            fn[k] = actn ? red[L.at(line, k, 2 * i + 2)] : id;
        }
        T A[2], B[2], C[2], D[4], ib[2];
        compact_rows<T>(f, l, fn, A, B, C, D, ib);
        pcr_solve<T>(A, B, C, D, i, Rl, RP);
        T xlp[4], xl[4], xf[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const T t = lane_prev(D[v]);
            xlp[v] = i > 0 ? t : (T)0;
            xl[v] = D[v];
        }
        first_unknowns<T>(f, ib, xlp, xl, xf);
        // the reads of row 2i+2 above (lane i) precede this write (lane i+1,
        // same wavefront) in program order
        if (act) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                red[L.at(line, 6 + v, 2 * i)] = xf[v];
                red[L.at(line, 6 + v, 2 * i + 1)] = xl[v];
            }
        }
    }
}

// Reduced system in 8 values per chunk, for kernels whose LDS is mostly taken
// by other data: the chunk's lane divides its first row by b and stores it;
// after a barrier it reads the next chunk's normalised first row, eliminates
// both first unknowns from its last row (the compaction step above) and stores
// that row, again divided by its b.  PCR then runs over one row per chunk, and
// each lane recovers x_f from its own normalised first row.  Row i of line
// `line` is at NormLayout::at; chunks are XOR-swizzled by the line so that the
// chunk-major stores (lanes = lines) and the line-major PCR loads (lanes =
// chunks) both spread over the banks.
constexpr int kNormFields = 8;  // a_v c_v a_t c_t d_u d_v d_w d_t, over the row's b
struct NormLayout {
    int P2;  // chunks per line rounded up to a power of two
    __device__ __forceinline__ int at(int line, int field, int chunk) const {
        return (line * kNormFields + field) * P2 + (chunk ^ (line & (P2 - 1)));
    }
};

__host__ __device__ constexpr int pow2_at_least(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

// First row (kRedFields layout) -> normalised: x_f = d - a x_l(prev) - c x_l.
template <class T>
__device__ __forceinline__ void normalise_first(const T* f, T* fh) {
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const T ib = rcp(f[3 * q + 1]);
        fh[2 * q] = f[3 * q] * ib;
        fh[2 * q + 1] = f[3 * q + 2] * ib;
#pragma unroll
        for (int v = (q == 0 ? 0 : 3); v < (q == 0 ? 3 : 4); v++) fh[4 + v] = f[6 + v] * ib;
    }
}

// Last row l with this chunk's (fh) and the next chunk's (nx, zeros past the
// line's end) normalised first rows substituted; normalised result in cr.
template <class T>
__device__ __forceinline__ void compact_norm(const T* fh, const T* l, const T* nx, T* cr) {
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const T al = l[3 * q], bl = l[3 * q + 1], cl = l[3 * q + 2];
        const T iB = rcp(bl - al * fh[2 * q + 1] - cl * nx[2 * q]);
        cr[2 * q] = -al * fh[2 * q] * iB;
        cr[2 * q + 1] = -cl * nx[2 * q + 1] * iB;
#pragma unroll
        for (int v = (q == 0 ? 0 : 3); v < (q == 0 ? 3 : 4); v++)
            cr[4 + v] = (l[6 + v] - al * fh[4 + v] - cl * nx[4 + v]) * iB;
    }
}

// PCR over the compacted rows of `nl` lines (P chunks each) with the lanes of
// `nw` wavefronts (this one is wave `w`), the lanes of a line inside one
// wavefront; x_l replaces the right-hand sides (fields 4..7) in place.
template <class T>
__device__ __forceinline__ void solve_norm_lines(T* red, NormLayout NL, int nl, int P, int w, int nw) {
    const int lane = threadIdx.x & 63;
    const int RP = NL.P2;
    const int lpp = 64 / RP;
    const int i = lane % RP;
    for (int base = w * lpp; base < nl; base += nw * lpp) {
        const int line = base + lane / RP;
        const bool act = line < nl && i < P;
        T A[2], B[2], C[2], D[4];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            A[q] = act ? red[NL.at(line, 2 * q, i)] : (T)0;
            B[q] = (T)1;
            C[q] = act ? red[NL.at(line, 2 * q + 1, i)] : (T)0;
        }
#pragma unroll
        for (int v = 0; v < 4; v++) D[v] = act ? red[NL.at(line, 4 + v, i)] : (T)0;
        pcr_solve<T>(A, B, C, D, i, line < nl ? P : 0, RP);
        if (act) {
#pragma unroll
            for (int v = 0; v < 4; v++) red[NL.at(line, 4 + v, i)] = D[v];
        }
    }
}

// 16-byte vector of consecutive elements.
template <class T>
struct Vec16;
template <>
struct Vec16<float> {
    using type = float4;
    static constexpr int n = 4;
};
template <>
struct Vec16<double> {
    using type = double2;
    static constexpr int n = 2;
};

// r[0..M) <- p[0..M): 16-byte loads when `vec` (all M rows inside the line and
// aligned), else element loads of the first `valid` rows (zero beyond).
template <class T, int M>
__device__ __forceinline__ void load_run(const T* __restrict__ p, T* r, bool vec, int valid) {
    constexpr int VW = Vec16<T>::n;
    if constexpr (M % VW == 0) {
        if (vec) {
            using V = typename Vec16<T>::type;
#pragma unroll
            for (int k = 0; k < M / VW; k++) {
                V v = reinterpret_cast<const V*>(p)[k];
                const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
                for (int q = 0; q < VW; q++) r[k * VW + q] = e[q];
            }
            return;
        }
    }
#pragma unroll
    for (int j = 0; j < M; j++) r[j] = j < valid ? p[j] : (T)0;
}

// r[0..M) <- bytes p[0..M): one M-byte load when `vec` and p is M-byte aligned
// (M = 2, 4, 8, 16), else byte loads of the first `valid` bytes (zero beyond).
template <int M>
__device__ __forceinline__ void load_bytes(const uint8_t* __restrict__ p, uint8_t* r, bool vec,
                                           int valid) {
    if (vec && ((uintptr_t)p % M) == 0) {
        if constexpr (M == 2) {
            const uint16_t w = *reinterpret_cast<const uint16_t*>(p);
            r[0] = (uint8_t)w; r[1] = (uint8_t)(w >> 8);
            return;
        } else if constexpr (M == 4 || M == 8 || M == 16) {
            using W = typename std::conditional<M == 4, uint32_t,
                      typename std::conditional<M == 8, uint2, uint4>::type>::type;
            const W w = *reinterpret_cast<const W*>(p);
            const uint8_t* b = reinterpret_cast<const uint8_t*>(&w);
#pragma unroll
            for (int j = 0; j < M; j++) r[j] = b[j];
            return;
        }
    }
#pragma unroll
    for (int j = 0; j < M; j++) r[j] = j < valid ? p[j] : (uint8_t)0;
}

// A store with a cache hint: NT streams it past the caches (non-temporal),
// for sweeps whose layers far exceed the memory-side cache, where written
// values would only evict values that the same sweep still has to read.
template <bool NT, class T>
__device__ __forceinline__ void put_hint(T* p, T v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <class T, int M, bool NT = false>
__device__ __forceinline__ void store_run(T* __restrict__ p, const T* r, bool vec, int valid) {
    constexpr int VW = Vec16<T>::n;
    if constexpr (M % VW == 0) {
        if (vec) {
            typedef T NV __attribute__((ext_vector_type(VW)));
#pragma unroll
            for (int k = 0; k < M / VW; k++) {
                NV v;
#pragma unroll
                for (int q = 0; q < VW; q++) v[q] = r[k * VW + q];
                put_hint<NT>(reinterpret_cast<NV*>(p) + k, v);
            }
            return;
        }
    }
#pragma unroll
    for (int j = 0; j < M; j++)
        if (j < valid) put_hint<NT>(p + j, r[j]);
}

// Runs f(std::true_type) when nt, else f(std::false_type): a uniform branch
// around a whole store phase.  The empty asm statements at the ends of the
// streaming side keep the optimiser from merging the two sides' identical
// stores into plain ones (the hint is metadata, dropped on merging).
template <class F>
__device__ __forceinline__ void with_store_hint(bool nt, F&& f) {
    if (nt) {
        asm volatile("");
        f(std::true_type{});
        asm volatile("");
    } else {
        f(std::false_type{});
    }
}

// Loads of one lane's run of M rows at c0 (the first `valid` of them inside
// the line), all issued whatever `valid` is: rows past it read from node 0
// (addresses masked, not selected, so that no load becomes conditional).
// `vall`: every run of the launch is whole and 16-byte aligned (vector loads).
template <class T, int M, bool VALL>
__device__ __forceinline__ void load_run_raw(const T* __restrict__ p, int64_t c0, int valid, T* r) {
    constexpr int VW = Vec16<T>::n;
    if constexpr (M % VW == 0 && VALL) {
        {
            using V = typename Vec16<T>::type;
            const V* q = reinterpret_cast<const V*>(p + (c0 & -(int64_t)(valid > 0)));
#pragma unroll
            for (int k = 0; k < M / VW; k++) {
                const V v = q[k];
                const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
                for (int i = 0; i < VW; i++) r[k * VW + i] = e[i];
            }
            return;
        }
    }
#pragma unroll
    for (int j = 0; j < M; j++) r[j] = p[(c0 + j) & -(int64_t)(j < valid)];
}

template <int M, bool VALL>
__device__ __forceinline__ void load_bytes_raw(const uint8_t* __restrict__ p, int64_t c0, int valid,
                                               int* r) {
    if constexpr ((M == 4 || M == 8) && VALL) {
        {
            const uint32_t* q = reinterpret_cast<const uint32_t*>(p + (c0 & -(int64_t)(valid > 0)));
#pragma unroll
            for (int k = 0; k < M / 4; k++) {
                const uint32_t v = q[k];
#pragma unroll
                for (int i = 0; i < 4; i++) r[4 * k + i] = (v >> (8 * i)) & 0xff;
            }
            return;
        }
    }
#pragma unroll
    for (int j = 0; j < M; j++) r[j] = p[(c0 + j) & -(int64_t)(j < valid)];
}

template <class T>
__device__ __forceinline__ void pin(const T& v) {
    asm volatile("" ::"v"(v));
}

}  // namespace lsolve
}  // namespace cmcx
