// Colatitude (Legendre) and radial contractions as MFMA GEMMs, FP64 or FP32.
//
// Replaces the hot loops
//   integrate_latitudinal  include/zernike_glq_transformer.hpp:778-817
//   integrate_radial       include/zernike_glq_transformer.hpp:819-865
//   sum_n                  include/zernike_glq_transformer.hpp:867-899
//   sum_l                  include/zernike_glq_transformer.hpp:901-941
//   integrate_latitudinal  include/sh_glq_transformer.hpp:886-974
//   sum_l                  include/sh_glq_transformer.hpp:976-1051
//
// Every stage is a set of independent small GEMMs ("problems"): one per
// (m, parity) for the Legendre stages -- the parity split
// P_lm(-z) = (-1)^(l+m) P_lm(z) halves both the contraction length and the
// flops relative to zest's dense sum over all latitudes -- and one per l for
// the radial stages. A workgroup of four waves owns a tile of up to 16*NS rows
// and 64 columns of one problem. The tile is cut into 16x16 MFMA blocks
// ((strip, column block) pairs) dealt round-robin to the waves, so a short
// problem pays for whole 16-row strips only.
//
// Operands are described by row base pointers: every operand is a set of
// contiguous rows, either one row per output index (contiguous along the
// contraction axis, "kind 0") or one row per contraction index (contiguous
// along the output axis, "kind 1"). Each thread loads fixed row segments, so
// the loaders need no division, and the next chunk of the contraction axis is
// fetched into registers while the current chunk runs on the MFMA units.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "zgl_internal.hpp"

namespace zgl {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <class T> struct Mfma;
template <> struct Mfma<double> {
    using acc_t = dbl4;
    __device__ static acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // v_mfma_f64_16x16x4_f64: lane holds D[lane/16 + 4r][lane%16]
    __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
    using acc_t = flt4;
    __device__ static acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // v_mfma_f32_16x16x4_f32: lane holds D[4(lane/16) + r][lane%16]
    __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int TN = 64;     // columns per tile
constexpr int TK = 32;     // contraction chunk
constexpr int kThreads = 256;
// the operand loaders map 32 consecutive threads onto one chunk row (and the
// pair loader 64 threads onto 32 (k, re/im) pairs)
static_assert(TK == 32 && TN == 64 && kThreads == 256, "loader thread mapping");

// Operand kinds (see the header comment).
enum { ROWS_OUT = 0, ROWS_K = 1 };

// Workgroup b runs on XCD b % 8 (each XCD has its own L2). The work items
// form groups of kGroupItems consecutive items (the row tiles of one problem,
// as the work lists are built), dealt to the XCDs in turn (neighbouring
// problems, of similar cost, on different XCDs); an XCD walks each of its
// groups column tile by column tile, the items of a group consecutively: the
// workgroups of one (group, column tile) then read the same B tile and run
// side by side, so it comes from HBM once, and the group's A blocks (P_lm or
// R_nl table blocks, re-staged for every column tile) stay in the L2 while
// the group's column tiles pass. (ZGL_GEMM_GROUP=n sets another group size,
// for comparison runs; 1: every item's column tiles in turn.)
constexpr int kGroupItems = 4;

int gemm_group() {
    static const int g = [] {
        const char* e = std::getenv("ZGL_GEMM_GROUP");
        const int v = e ? std::atoi(e) : kGroupItems;
        return v >= 1 && v <= 64 ? v : kGroupItems;
    }();
    return g;
}

template <class S, int NS>
__global__ __launch_bounds__(kThreads) void gemm_kernel(S st, const int4* __restrict__ work, int nwork,
                                                        int nct, int gsz) {
    using T = typename S::T;
    using M = Mfma<T>;
    constexpr int TM = 16 * NS;
    const int xcd = blockIdx.x & 7, q = blockIdx.x >> 3;
    const int gl = q / (nct * gsz), rq = q - gl * (nct * gsz);
    const int ct = rq / gsz, r = rq - ct * gsz;
    const int wi = (gl * 8 + xcd) * gsz + r;  // group gl * 8 + xcd, its item r
    if (wi >= nwork) return;
    const int4 wk = work[wi];
    const int n0 = ct * TN;
    const typename S::Prob P = st.problem(wk.x, blockIdx.z);
    const int m0 = wk.y;
    if (!P.active || m0 >= P.M || n0 >= P.N) return;

    __shared__ T As[TM][TK + 1];
    __shared__ T Bs[TK][TN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rows = min(TM, P.M - m0), cols = min(TN, P.N - n0);
    const int strips = (rows + 15) >> 4, blocks = (cols + 15) >> 4;
    const int npairs = strips * blocks;

    // A tile TM x TK. ROWS_OUT: thread -> (i = tid / 32 + 8q, k = tid % 32);
    // ROWS_K: thread -> (k = tid / 32 + 8 (q % 4), i = tid % 32 + 32 (q / 4)).
    // Consecutive threads always read consecutive elements of one row.
    constexpr int A_NQ = (S::A_KIND == ROWS_OUT) ? (TM + 7) / 8 : (TK / 8) * ((TM + 31) / 32);
    constexpr int B_NQ = (S::B_KIND == ROWS_OUT) ? (TN / 8) : (TK / 4);
    // two register buffers per operand: the chunk after next is in flight while
    // the current one is computed (loads get two chunks of MFMA work to land)
    T ra0[A_NQ], rb0[B_NQ], ra1[A_NQ], rb1[B_NQ];

    auto load_a = [&](int k0, T (&ra)[A_NQ]) {
        if constexpr (S::A_KIND == ROWS_OUT) {
            const int k = tid & (TK - 1);
#pragma unroll
            for (int q = 0; q < A_NQ; ++q) {
                const int i = (tid >> 5) + 8 * q;
                ra[q] = (i < rows && k0 + k < P.K) ? (T)st.a_row(P, m0 + i)[k0 + k] : T(0);
            }
        } else {
            const int il = tid & 31;
#pragma unroll
            for (int q = 0; q < A_NQ; ++q) {
                const int k = (tid >> 5) + 8 * (q % (TK / 8));
                const int i = il + 32 * (q / (TK / 8));
                ra[q] = (i < rows && k0 + k < P.K) ? (T)st.a_row(P, k0 + k)[m0 + i] : T(0);
            }
        }
    };
    auto store_a = [&](const T (&ra)[A_NQ]) {
        if constexpr (S::A_KIND == ROWS_OUT) {
            const int k = tid & (TK - 1);
#pragma unroll
            for (int q = 0; q < A_NQ; ++q) {
                const int i = (tid >> 5) + 8 * q;
                if (i < TM) As[i][k] = ra[q];
            }
        } else {
            const int il = tid & 31;
#pragma unroll
            for (int q = 0; q < A_NQ; ++q) {
                const int k = (tid >> 5) + 8 * (q % (TK / 8));
                const int i = il + 32 * (q / (TK / 8));
                if (i < TM) As[i][k] = ra[q];
            }
        }
    };
    // B: ROWS_K -> thread (k = tid / 64 + 4q, j = tid % 64); ROWS_OUT -> rows j
    // hold the contraction axis: thread (j = tid / 32 + 8q, k = tid % 32); with
    // B_PAIR the rows are indexed by j / 2 and hold (k, re/im) pairs.
    auto load_b = [&](int k0, T (&rb)[B_NQ]) {
        if constexpr (S::B_KIND == ROWS_K) {
            const int j = tid & 63;
#pragma unroll
            for (int q = 0; q < B_NQ; ++q) {
                const int k = (tid >> 6) + 4 * q;
                rb[q] = (k0 + k < P.K && j < cols) ? (T)st.b_row(P, k0 + k)[n0 + j] : T(0);
            }
        } else if constexpr (!S::B_PAIR) {
            const int k = tid & (TK - 1);
#pragma unroll
            for (int q = 0; q < B_NQ; ++q) {
                const int j = (tid >> 5) + 8 * q;
                rb[q] = (k0 + k < P.K && j < cols) ? (T)st.b_row(P, n0 + j)[k0 + k] : T(0);
            }
        } else {
            // pairs: thread -> (jj = tid / 64 + 4q pair row, e = tid % 64 over (k, ri))
            const int e = tid & 63, k = e >> 1, ri = e & 1;
#pragma unroll
            for (int q = 0; q < B_NQ; ++q) {
                const int jj = (tid >> 6) + 4 * q;
                const int j = 2 * jj + ri;
                rb[q] = (k0 + k < P.K && j < cols) ? (T)st.b_row(P, (n0 >> 1) + jj)[2 * (k0 + k) + ri] : T(0);
            }
        }
    };
    auto store_b = [&](const T (&rb)[B_NQ]) {
        if constexpr (S::B_KIND == ROWS_K) {
            const int j = tid & 63;
#pragma unroll
            for (int q = 0; q < B_NQ; ++q) Bs[(tid >> 6) + 4 * q][j] = rb[q];
        } else if constexpr (!S::B_PAIR) {
            const int k = tid & (TK - 1);
#pragma unroll
            for (int q = 0; q < B_NQ; ++q) Bs[k][(tid >> 5) + 8 * q] = rb[q];
        } else {
            const int e = tid & 63, k = e >> 1, ri = e & 1;
#pragma unroll
            for (int q = 0; q < B_NQ; ++q) Bs[k][2 * ((tid >> 6) + 4 * q) + ri] = rb[q];
        }
    };

    typename M::acc_t acc[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) acc[q] = typename M::acc_t{0, 0, 0, 0};
    // FP32 products: each chunk's FP32 partial sums are folded into FP64
    // accumulators (the contraction runs over hundreds of latitudes or degrees
    // at high order; within a 32-deep chunk the FP32 sum is short)
    constexpr bool kFold = std::is_same_v<T, float>;
    dbl4 acc64[kFold ? NS : 1];
#pragma unroll
    for (int q = 0; q < (kFold ? NS : 1); ++q) acc64[q] = dbl4{0, 0, 0, 0};
    auto fold = [&] {
        if constexpr (kFold) {
#pragma unroll
            for (int q = 0; q < NS; ++q) {
#pragma unroll
                for (int r = 0; r < 4; ++r) acc64[q][r] += (double)acc[q][r];
                acc[q] = typename M::acc_t{0, 0, 0, 0};
            }
        }
    };

    auto compute = [&](int k0) {
        const int kpad = (min(TK, P.K - k0) + 3) & ~3;
        for (int kk = 0; kk < kpad; kk += 4) {
            const int kr = kk + (lane >> 4);
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                const int p = wave + 4 * q;
                if (p < npairs) {
                    const int s = p / blocks, b = p - s * blocks;
                    acc[q] = M::mma(As[s * 16 + (lane & 15)][kr], Bs[kr][b * 16 + (lane & 15)], acc[q]);
                }
            }
        }
    };
    // one chunk: stage its registers, refill them with the chunk two ahead,
    // compute from LDS
    auto chunk = [&](int k0, T (&ra)[A_NQ], T (&rb)[B_NQ]) {
        store_a(ra);
        store_b(rb);
        __syncthreads();
        if (k0 + 2 * TK < P.K) {
            load_a(k0 + 2 * TK, ra);
            load_b(k0 + 2 * TK, rb);
        }
        compute(k0);
        fold();
        __syncthreads();
    };

    load_a(0, ra0);
    load_b(0, rb0);
    if (TK < P.K) {
        load_a(TK, ra1);
        load_b(TK, rb1);
    }
    for (int k0 = 0; k0 < P.K; k0 += 2 * TK) {
        chunk(k0, ra0, rb0);
        if (k0 + TK < P.K) chunk(k0 + TK, ra1, rb1);
    }
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        const int p = wave + 4 * q;
        if (p >= npairs) continue;
        const int s = p / blocks, b = p - s * blocks;
        const int j = b * 16 + (lane & 15);
        if (j >= cols) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = s * 16 + M::row(lane, r);
            if constexpr (kFold) {
                if (i < rows) st.c(P, m0 + i, n0 + j, (T)acc64[q][r]);
            } else {
                if (i < rows) st.c(P, m0 + i, n0 + j, acc[q][r]);
            }
        }
    }
}

// ---------------------------------------------------------------- Zernike
// Legendre forward: flm[l,m][k] = sum_h P_lm(z_h) X^par_m[h][k],
// l = m + par + 2i (rows i), columns (k, re/im).
template <class TT>
struct ZtLegendreFwd {
    using T = TT;
    static constexpr int A_KIND = ROWS_OUT, B_KIND = ROWS_K;
    static constexpr bool B_PAIR = false;
    const T* ptab; const T* spec; T* flm;
    long spec_stride, flm_stride;  // elements of T
    int order, nt, jh, ncol;
    struct Prob { int active, M, N, K, m, par; const T* A; const T* B; T* C; };
    __device__ Prob problem(int prob, int inst) const {
        Prob p;
        p.m = prob >> 1; p.par = prob & 1;
        p.M = plm_count(nt, p.m, p.par); p.K = jh; p.N = 2 * ncol;
        p.active = p.M > 0;
        p.A = ptab + plm_offset(order, p.m, p.par, jh);
        p.B = spec + inst * spec_stride + (long)(p.m * 2 + p.par) * jh * 2 * ncol;
        p.C = flm + inst * flm_stride;
        return p;
    }
    __device__ const T* a_row(const Prob& p, int i) const { return p.A + (long)i * jh; }
    __device__ const T* b_row(const Prob& p, int k) const { return p.B + (long)k * 2 * ncol; }
    __device__ void c(const Prob& p, int i, int j, T v) const {
        const int l = p.m + p.par + 2 * i;
        p.C[tri_idx(l, p.m) * 2 * ncol + j] = v;
    }
};

// Legendre backward: X^par_m[h][k] = sum_i P_lm(z_h) flm[l,m][k].
template <class TT>
struct ZtLegendreBwd {
    using T = TT;
    static constexpr int A_KIND = ROWS_K, B_KIND = ROWS_K;
    static constexpr bool B_PAIR = false;
    const T* ptab; const T* flm; T* spec;
    long spec_stride, flm_stride;
    int order, nt, jh, ncol;
    struct Prob { int active, M, N, K, m, par; const T* A; const T* B; T* C; };
    __device__ Prob problem(int prob, int inst) const {
        Prob p;
        p.m = prob >> 1; p.par = prob & 1;
        p.M = jh; p.K = plm_count(nt, p.m, p.par); p.N = 2 * ncol;
        p.active = p.m < nt;
        p.A = ptab + plm_offset(order, p.m, p.par, jh);
        p.B = flm + inst * flm_stride;
        p.C = spec + inst * spec_stride + (long)(p.m * 2 + p.par) * jh * 2 * ncol;
        return p;
    }
    __device__ const T* a_row(const Prob& p, int k) const { return p.A + (long)k * jh; }
    __device__ const T* b_row(const Prob& p, int k) const {
        return p.B + tri_idx(p.m + p.par + 2 * k, p.m) * 2 * ncol;
    }
    __device__ void c(const Prob& p, int i, int j, T v) const { p.C[(long)i * 2 * ncol + j] = v; }
};

// Radial forward: c[n,l,m] = kappa_n sum_k R_nl(r_k) flm[l,m][k], n = l + 2i.
template <class TT>
struct ZtRadialFwd {
    using T = TT;
    static constexpr int A_KIND = ROWS_OUT, B_KIND = ROWS_OUT;
    static constexpr bool B_PAIR = true;
    const T* rtab; const T* flm; Inst<double> coeffs;
    long flm_stride;  // elements of T
    int nt, nr, unnormed;
    struct Prob { int active, M, N, K, l; const T* B; double* C; };
    __device__ Prob problem(int prob, int inst) const {
        Prob p;
        p.l = prob;
        p.M = (nt - p.l + 1) / 2; p.K = nr; p.N = 2 * (p.l + 1);
        p.active = p.l < nt;
        p.B = flm + inst * flm_stride + tri_idx(p.l, 0) * 2 * nr;
        p.C = coeffs.at(inst);
        return p;
    }
    __device__ const T* a_row(const Prob& p, int i) const {
        return rtab + rad_idx(p.l + 2 * i, p.l) * nr;
    }
    __device__ const T* b_row(const Prob& p, int m) const { return p.B + (long)m * 2 * nr; }
    __device__ void c(const Prob& p, int i, int j, T v) const {
        const int n = p.l + 2 * i;
        const double kappa = unnormed ? (double)(2 * n + 3) : 1.0;
        p.C[tet_idx(n, p.l, 0) * 2 + j] = (double)v * kappa;
    }
};

// Radial backward: flm[l,m][k] = sum_i R_nl(r_k) c[n,l,m], n = l + 2i;
// rows (m, re/im), columns k.
template <class TT>
struct ZtRadialBwd {
    using T = TT;
    static constexpr int A_KIND = ROWS_K, B_KIND = ROWS_K;
    static constexpr bool B_PAIR = false;
    const T* rtab; Inst<const double> coeffs; T* flm;
    long flm_stride;
    int nt, nr;
    struct Prob { int active, M, N, K, l; const double* A; T* C; };
    __device__ Prob problem(int prob, int inst) const {
        Prob p;
        p.l = prob;
        p.M = 2 * (p.l + 1); p.K = (nt - p.l + 1) / 2; p.N = nr;
        p.active = p.l < nt;
        p.A = coeffs.at(inst);
        p.C = flm + inst * flm_stride + tri_idx(p.l, 0) * 2 * nr;
        return p;
    }
    __device__ const double* a_row(const Prob& p, int k) const {
        return p.A + tet_idx(p.l + 2 * k, p.l, 0) * 2;
    }
    __device__ const T* b_row(const Prob& p, int k) const {
        return rtab + rad_idx(p.l + 2 * k, p.l) * nr;
    }
    __device__ void c(const Prob& p, int i, int j, T v) const {
        p.C[(long)(i >> 1) * 2 * nr + 2 * j + (i & 1)] = v;
    }
};

// ------------------------------------------------------ spherical harmonics
// The instances of a chunk form the column axis: columns (instance, re/im).
// The coefficients go through a coefficient-major staging array
// stage[tri_idx(l, m)][instance][re/im] (contiguous along the column axis);
// st_stage_coeffs moves them to and from the per-instance arrays with
// coalesced accesses on both sides.
template <class TT>
struct StLegendreFwd {
    using T = TT;
    static constexpr int A_KIND = ROWS_OUT, B_KIND = ROWS_OUT;
    static constexpr bool B_PAIR = true;
    const T* ptab; const T* spec; T* stage;
    long spec_stride;  // elements of T
    int order, nt, jh, nb;
    struct Prob { int active, M, N, K, m, par; const T* A; long boff; };
    __device__ Prob problem(int prob, int) const {
        Prob p;
        p.m = prob >> 1; p.par = prob & 1;
        p.M = plm_count(nt, p.m, p.par); p.K = jh; p.N = 2 * nb;
        p.active = p.M > 0;
        p.A = ptab + plm_offset(order, p.m, p.par, jh);
        p.boff = (long)(p.m * 2 + p.par) * jh * 2;
        return p;
    }
    __device__ const T* a_row(const Prob& p, int i) const { return p.A + (long)i * jh; }
    __device__ const T* b_row(const Prob& p, int b) const { return spec + (long)b * spec_stride + p.boff; }
    __device__ void c(const Prob& p, int i, int j, T v) const {
        const int l = p.m + p.par + 2 * i;
        stage[(long)tri_idx(l, p.m) * (2 * nb) + j] = v;
    }
};

// Backward: the coefficient operand is the staging array (rows l of the
// problem's m, columns (instance, re/im)).
template <class TT>
struct StLegendreBwd {
    using T = TT;
    static constexpr int A_KIND = ROWS_K, B_KIND = ROWS_K;
    static constexpr bool B_PAIR = false;
    const T* ptab; const T* stage; T* spec;
    long spec_stride;
    int order, nt, jh, nb;
    struct Prob { int active, M, N, K, m, par; const T* A; long soff; };
    __device__ Prob problem(int prob, int) const {
        Prob p;
        p.m = prob >> 1; p.par = prob & 1;
        p.M = jh; p.K = plm_count(nt, p.m, p.par); p.N = 2 * nb;
        p.active = p.m < nt;
        p.A = ptab + plm_offset(order, p.m, p.par, jh);
        p.soff = (long)(p.m * 2 + p.par) * jh * 2;
        return p;
    }
    __device__ const T* a_row(const Prob& p, int k) const { return p.A + (long)k * jh; }
    __device__ const T* b_row(const Prob& p, int k) const {
        return stage + (long)tri_idx(p.m + p.par + 2 * k, p.m) * (2 * nb);
    }
    __device__ void c(const Prob& p, int i, int j, T v) const {
        spec[(long)(j >> 1) * spec_stride + p.soff + 2 * i + (j & 1)] = v;
    }
};

// Coefficients <-> staging: a 32 x 32 tile of (coefficient, instance) pairs
// through LDS; each instance's array is read or written as 32 consecutive
// pairs, each staging row likewise. TO_COEFF: stage -> coefficients.
template <class T, bool TO_COEFF>
__global__ __launch_bounds__(kThreads) void stage_kernel(T* stage, Inst<double> coeffs, int ntri, int nb) {
    __shared__ double2 tile[32][33];
    const int t0 = blockIdx.x * 32, b0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const long row = 2L * nb;
    if constexpr (TO_COEFF) {
        for (int i = ty; i < 32; i += kThreads / 32) {
            const int t = t0 + i, b = b0 + tx;
            if (t < ntri && b < nb) {
                const T* sp = stage + (long)t * row + 2 * b;
                tile[i][tx] = make_double2((double)sp[0], (double)sp[1]);
            }
        }
        __syncthreads();
        for (int i = ty; i < 32; i += kThreads / 32) {
            const int b = b0 + i, t = t0 + tx;
            if (t < ntri && b < nb) {
                const double2 v = tile[tx][i];
                double* c = coeffs.at(b) + 2L * t;
                c[0] = v.x;
                c[1] = v.y;
            }
        }
    } else {
        for (int i = ty; i < 32; i += kThreads / 32) {
            const int b = b0 + i, t = t0 + tx;
            if (t < ntri && b < nb) {
                const double* c = coeffs.at(b) + 2L * t;
                tile[tx][i] = make_double2(c[0], c[1]);
            }
        }
        __syncthreads();
        for (int i = ty; i < 32; i += kThreads / 32) {
            const int t = t0 + i, b = b0 + tx;
            if (t < ntri && b < nb) {
                const double2 v = tile[i][tx];
                T* sp = stage + (long)t * row + 2 * b;
                sp[0] = (T)v.x;
                sp[1] = (T)v.y;
            }
        }
    }
}

template <class T, bool TO_COEFF>
void st_stage_coeffs(T* stage, Inst<double> coeffs, int nt, int nb, hipStream_t s) {
    const int ntri = nt * (nt + 1) / 2;
    if (ntri == 0 || nb <= 0) return;
    dim3 grid((ntri + 31) / 32, (nb + 31) / 32);
    stage_kernel<T, TO_COEFF><<<grid, kThreads, 0, s>>>(stage, coeffs, ntri, nb);
    check_hip(hipGetLastError(), "coefficient staging launch");
}

template <int NS, class S>
void launch(const S& st, const int4* work, int nwork, int ncols, int nz, hipStream_t s) {
    if (nwork == 0 || ncols <= 0 || nz <= 0) return;
    const int nct = (ncols + TN - 1) / TN;
    const int gsz = gemm_group();
    const int groups = (nwork + gsz - 1) / gsz;  // dealt to the 8 XCDs in turn
    dim3 grid((unsigned)(8 * nct * gsz) * (unsigned)((groups + 7) / 8), 1, nz);
    gemm_kernel<S, NS><<<grid, kThreads, 0, s>>>(st, work, nwork, nct, gsz);
    check_hip(hipGetLastError(), "gemm launch");
}

template <class T>
const T* ptab_of(const OrderTables& t);
template <> const double* ptab_of<double>(const OrderTables& t) { return t.ptab; }
template <> const float* ptab_of<float>(const OrderTables& t) { return t.ptab32; }
template <class T>
const T* rtab_of(const OrderTables& t);
template <> const double* rtab_of<double>(const OrderTables& t) { return t.rtab; }
template <> const float* rtab_of<float>(const OrderTables& t) { return t.rtab32; }

}  // namespace

bool legendre_stream_enabled() {
    const char* e = std::getenv("ZGL_LEGENDRE_ENGINE");
    return !(e && std::strcmp(e, "tiled") == 0);
}

template <class T>
void zt_legendre_forward(const ZtGemmArgs& a, hipStream_t s) {
    if ((legendre_stream_enabled() || a.t->plm_in_kernel) && zt_legendre_forward_stream<T>(a, s)) return;
    if (a.t->plm_in_kernel) throw Error{ZGL_ERROR_NOT_SUPPORTED, "P_lm recomputation needs the streaming kernels"};
    const Geom& g = a.t->g;
    ZtLegendreFwd<T> st{ptab_of<T>(*a.t), (const T*)a.spec, (T*)a.flm, a.spec_stride * 2,
                        a.flm_stride * 2, g.order, a.nt, g.jh, g.ncol};
    launch<kNsLegendreFwd>(st, a.t->work_lf, a.t->n_lf, 2 * g.ncol, a.nb, s);
}

template <class T>
void zt_legendre_backward(const ZtGemmArgs& a, hipStream_t s) {
    if ((legendre_stream_enabled() || a.t->plm_in_kernel) && zt_legendre_backward_stream<T>(a, s)) return;
    if (a.t->plm_in_kernel) throw Error{ZGL_ERROR_NOT_SUPPORTED, "P_lm recomputation needs the streaming kernels"};
    const Geom& g = a.t->g;
    ZtLegendreBwd<T> st{ptab_of<T>(*a.t), (const T*)a.flm, (T*)a.spec, a.spec_stride * 2,
                        a.flm_stride * 2, g.order, a.nt, g.jh, g.ncol};
    launch<kNsLegendreBwdZt>(st, a.t->work_lb, a.t->n_lb, 2 * g.ncol, a.nb, s);
}

template <class T>
void zt_radial_forward(const ZtGemmArgs& a, hipStream_t s) {
    if (legendre_stream_enabled() && zt_radial_forward_stream<T>(a, s)) return;
    const Geom& g = a.t->g;
    ZtRadialFwd<T> st{rtab_of<T>(*a.t), (const T*)a.flm, a.coeffs, a.flm_stride * 2,
                      a.nt, g.ncol, a.zernike_unnormed};
    launch<kNsRadial>(st, a.t->work_rf, a.t->n_rf, 2 * g.order, a.nb, s);
}

template <class T>
void zt_radial_backward(const ZtGemmArgs& a, hipStream_t s) {
    if (legendre_stream_enabled() && zt_radial_backward_stream<T>(a, s)) return;
    const Geom& g = a.t->g;
    ZtRadialBwd<T> st{rtab_of<T>(*a.t), a.coeffs, (T*)a.flm, a.flm_stride * 2, a.nt,
                      g.ncol};
    launch<kNsRadial>(st, a.t->work_rb, a.t->n_rb, g.ncol, a.nb, s);
}

template <class T>
void st_legendre_forward(const StGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    StLegendreFwd<T> st{ptab_of<T>(*a.t), (const T*)a.spec, (T*)a.stage, a.spec_stride * 2,
                        g.order, a.nt, g.jh, a.nb};
    launch<kNsLegendreFwd>(st, a.t->work_lf, a.t->n_lf, 2 * a.nb, 1, s);
    st_stage_coeffs<T, true>((T*)a.stage, a.coeffs, a.nt, a.nb, s);
}

template <class T>
void st_legendre_backward(const StGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    st_stage_coeffs<T, false>((T*)a.stage, a.coeffs, a.nt, a.nb, s);
    StLegendreBwd<T> st{ptab_of<T>(*a.t), (const T*)a.stage, (T*)a.spec, a.spec_stride * 2,
                        g.order, a.nt, g.jh, a.nb};
    launch<kNsLegendreBwdSt>(st, a.t->work_lb, a.t->n_lb, 2 * a.nb, 1, s);
}

#define ZGL_INSTANTIATE(T)                                                  \
    template void zt_legendre_forward<T>(const ZtGemmArgs&, hipStream_t);  \
    template void zt_legendre_backward<T>(const ZtGemmArgs&, hipStream_t); \
    template void zt_radial_forward<T>(const ZtGemmArgs&, hipStream_t);    \
    template void zt_radial_backward<T>(const ZtGemmArgs&, hipStream_t);   \
    template void st_legendre_forward<T>(const StGemmArgs&, hipStream_t);  \
    template void st_legendre_backward<T>(const StGemmArgs&, hipStream_t);
ZGL_INSTANTIATE(double)
ZGL_INSTANTIATE(float)

}  // namespace zgl
