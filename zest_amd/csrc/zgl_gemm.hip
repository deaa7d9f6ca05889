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
// problem pays for whole 16-row strips only, and the contraction axis is
// staged through LDS in chunks of 32 whose last chunk runs only the 4-deep
// MFMA steps it needs. Work lists of (problem, first row) are built per
// order on the host; grid.y walks the column tiles and grid.z the instances.
#include <hip/hip_runtime.h>

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

constexpr int TN = 64, TK = 32;

enum LoadOrder { K_FAST = 0, MN_FAST = 1, PAIR_K_FAST = 2 };

template <class S, int NS>
__global__ __launch_bounds__(256) void gemm_kernel(S st, const int4* __restrict__ work) {
    using T = typename S::T;
    using M = Mfma<T>;
    constexpr int TM = 16 * NS;
    const int4 wk = work[blockIdx.x];
    const int n0 = blockIdx.y * TN;
    const typename S::Prob P = st.problem(wk.x, blockIdx.z);
    const int m0 = wk.y;
    if (!P.active || m0 >= P.M || n0 >= P.N) return;

    __shared__ T As[TM][TK + 1];
    __shared__ T Bs[TK][TN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rows = min(TM, P.M - m0), cols = min(TN, P.N - n0);
    const int strips = (rows + 15) >> 4, blocks = (cols + 15) >> 4;
    const int npairs = strips * blocks;
    const int arows = strips * 16, bcols = blocks * 16;
    typename M::acc_t acc[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) acc[q] = typename M::acc_t{0, 0, 0, 0};

    for (int k0 = 0; k0 < P.K; k0 += TK) {
        const int kc = min(TK, P.K - k0);
        const int kpad = (kc + 3) & ~3;
        for (int e = tid; e < arows * kpad; e += 256) {
            int i, k;
            if (S::A_ORDER == K_FAST) { k = e % kpad; i = e / kpad; }
            else { i = e % arows; k = e / arows; }
            As[i][k] = (i < rows && k < kc) ? st.a(P, m0 + i, k0 + k) : T(0);
        }
        for (int e = tid; e < kpad * bcols; e += 256) {
            int k, j;
            if (S::B_ORDER == MN_FAST) { j = e % bcols; k = e / bcols; }
            else if (S::B_ORDER == K_FAST) { k = e % kpad; j = e / kpad; }
            else { const int ri = e & 1; k = (e >> 1) % kpad; j = 2 * ((e >> 1) / kpad) + ri; }
            Bs[k][j] = (k < kc && j < cols) ? st.b(P, k0 + k, n0 + j) : T(0);
        }
        __syncthreads();
        for (int kk = 0; kk < kpad; kk += 4) {
            const int kr = kk + (lane >> 4);
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                const int p = wave + 4 * q;
                if (p < npairs) {
                    const int s = p / blocks, b = p - s * blocks;
                    const T av = As[s * 16 + (lane & 15)][kr];
                    const T bv = Bs[kr][b * 16 + (lane & 15)];
                    acc[q] = M::mma(av, bv, acc[q]);
                }
            }
        }
        __syncthreads();
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
            if (i < rows) st.c(P, m0 + i, n0 + j, acc[q][r]);
        }
    }
}

// ---------------------------------------------------------------- Zernike
// Legendre forward: flm[l,m][k] = sum_h P_lm(z_h) X^par_m[h][k],
// l = m + par + 2i (rows i), columns (k, re/im).
template <class TT>
struct ZtLegendreFwd {
    using T = TT;
    static constexpr int A_ORDER = K_FAST, B_ORDER = MN_FAST;
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
    __device__ T a(const Prob& p, int i, int k) const { return p.A[(long)i * jh + k]; }
    __device__ T b(const Prob& p, int k, int j) const { return p.B[(long)k * 2 * ncol + j]; }
    __device__ void c(const Prob& p, int i, int j, T v) const {
        const int l = p.m + p.par + 2 * i;
        p.C[tri_idx(l, p.m) * 2 * ncol + j] = v;
    }
};

// Legendre backward: X^par_m[h][k] = sum_i P_lm(z_h) flm[l,m][k].
template <class TT>
struct ZtLegendreBwd {
    using T = TT;
    static constexpr int A_ORDER = MN_FAST, B_ORDER = MN_FAST;
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
    __device__ T a(const Prob& p, int i, int k) const { return p.A[(long)k * jh + i]; }
    __device__ T b(const Prob& p, int k, int j) const {
        const int l = p.m + p.par + 2 * k;
        return p.B[tri_idx(l, p.m) * 2 * ncol + j];
    }
    __device__ void c(const Prob& p, int i, int j, T v) const { p.C[(long)i * 2 * ncol + j] = v; }
};

// Radial forward: c[n,l,m] = kappa_n sum_k R_nl(r_k) flm[l,m][k], n = l + 2i.
template <class TT>
struct ZtRadialFwd {
    using T = TT;
    static constexpr int A_ORDER = K_FAST, B_ORDER = PAIR_K_FAST;
    const T* rtab; const T* flm; double* coeffs;
    long flm_stride, coeff_stride;  // elements of T, doubles
    int nt, nr, unnormed;
    struct Prob { int active, M, N, K, l; const T* B; double* C; };
    __device__ Prob problem(int prob, int inst) const {
        Prob p;
        p.l = prob;
        p.M = (nt - p.l + 1) / 2; p.K = nr; p.N = 2 * (p.l + 1);
        p.active = p.l < nt;
        p.B = flm + inst * flm_stride + tri_idx(p.l, 0) * 2 * nr;
        p.C = coeffs + inst * coeff_stride;
        return p;
    }
    __device__ T a(const Prob& p, int i, int k) const {
        return rtab[rad_idx(p.l + 2 * i, p.l) * nr + k];
    }
    __device__ T b(const Prob& p, int k, int j) const {
        return p.B[(long)(j >> 1) * 2 * nr + 2 * k + (j & 1)];
    }
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
    static constexpr int A_ORDER = MN_FAST, B_ORDER = MN_FAST;
    const T* rtab; const double* coeffs; T* flm;
    long flm_stride, coeff_stride;
    int nt, nr;
    struct Prob { int active, M, N, K, l; const double* A; T* C; };
    __device__ Prob problem(int prob, int inst) const {
        Prob p;
        p.l = prob;
        p.M = 2 * (p.l + 1); p.K = (nt - p.l + 1) / 2; p.N = nr;
        p.active = p.l < nt;
        p.A = coeffs + inst * coeff_stride;
        p.C = flm + inst * flm_stride + tri_idx(p.l, 0) * 2 * nr;
        return p;
    }
    __device__ T a(const Prob& p, int i, int k) const {
        return (T)p.A[tet_idx(p.l + 2 * k, p.l, 0) * 2 + i];
    }
    __device__ T b(const Prob& p, int k, int j) const {
        return rtab[rad_idx(p.l + 2 * k, p.l) * nr + j];
    }
    __device__ void c(const Prob& p, int i, int j, T v) const {
        p.C[(long)(i >> 1) * 2 * nr + 2 * j + (i & 1)] = v;
    }
};

// ------------------------------------------------------ spherical harmonics
// The instances of a chunk form the column axis: columns (instance, re/im).
template <class TT>
struct StLegendreFwd {
    using T = TT;
    static constexpr int A_ORDER = K_FAST, B_ORDER = PAIR_K_FAST;
    const T* ptab; const T* spec; double* coeffs;
    long spec_stride, coeff_stride;  // elements of T, doubles
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
    __device__ T a(const Prob& p, int i, int k) const { return p.A[(long)i * jh + k]; }
    __device__ T b(const Prob& p, int k, int j) const {
        return spec[(long)(j >> 1) * spec_stride + p.boff + 2 * k + (j & 1)];
    }
    __device__ void c(const Prob& p, int i, int j, T v) const {
        const int l = p.m + p.par + 2 * i;
        coeffs[(long)(j >> 1) * coeff_stride + tri_idx(l, p.m) * 2 + (j & 1)] = (double)v;
    }
};

template <class TT>
struct StLegendreBwd {
    using T = TT;
    static constexpr int A_ORDER = MN_FAST, B_ORDER = MN_FAST;
    const T* ptab; const double* coeffs; T* spec;
    long spec_stride, coeff_stride;
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
    __device__ T a(const Prob& p, int i, int k) const { return p.A[(long)k * jh + i]; }
    __device__ T b(const Prob& p, int k, int j) const {
        const int l = p.m + p.par + 2 * k;
        return (T)coeffs[(long)(j >> 1) * coeff_stride + tri_idx(l, p.m) * 2 + (j & 1)];
    }
    __device__ void c(const Prob& p, int i, int j, T v) const {
        spec[(long)(j >> 1) * spec_stride + p.soff + 2 * i + (j & 1)] = v;
    }
};

template <int NS, class S>
void launch(const S& st, const int4* work, int nwork, int ncols, int nz, hipStream_t s) {
    if (nwork == 0 || ncols <= 0 || nz <= 0) return;
    dim3 grid(nwork, (ncols + TN - 1) / TN, nz);
    gemm_kernel<S, NS><<<grid, 256, 0, s>>>(st, work);
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

template <class T>
void zt_legendre_forward(const ZtGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    ZtLegendreFwd<T> st{ptab_of<T>(*a.t), (const T*)a.spec, (T*)a.flm, a.spec_stride * 2,
                        a.flm_stride * 2, g.order, a.nt, g.jh, g.ncol};
    launch<kNsLegendreFwd>(st, a.t->work_lf, a.t->n_lf, 2 * g.ncol, a.nb, s);
}

template <class T>
void zt_legendre_backward(const ZtGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    ZtLegendreBwd<T> st{ptab_of<T>(*a.t), (const T*)a.flm, (T*)a.spec, a.spec_stride * 2,
                        a.flm_stride * 2, g.order, a.nt, g.jh, g.ncol};
    launch<kNsLegendreBwdZt>(st, a.t->work_lb, a.t->n_lb, 2 * g.ncol, a.nb, s);
}

template <class T>
void zt_radial_forward(const ZtGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    ZtRadialFwd<T> st{rtab_of<T>(*a.t), (const T*)a.flm, a.coeffs, a.flm_stride * 2, a.coeff_stride,
                      a.nt, g.ncol, a.zernike_unnormed};
    launch<kNsRadial>(st, a.t->work_rf, a.t->n_rf, 2 * g.order, a.nb, s);
}

template <class T>
void zt_radial_backward(const ZtGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    ZtRadialBwd<T> st{rtab_of<T>(*a.t), a.coeffs, (T*)a.flm, a.flm_stride * 2, a.coeff_stride, a.nt,
                      g.ncol};
    launch<kNsRadial>(st, a.t->work_rb, a.t->n_rb, g.ncol, a.nb, s);
}

template <class T>
void st_legendre_forward(const StGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    StLegendreFwd<T> st{ptab_of<T>(*a.t), (const T*)a.spec, a.coeffs, a.spec_stride * 2,
                        a.coeff_stride, g.order, a.nt, g.jh, a.nb};
    launch<kNsLegendreFwd>(st, a.t->work_lf, a.t->n_lf, 2 * a.nb, 1, s);
}

template <class T>
void st_legendre_backward(const StGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    StLegendreBwd<T> st{ptab_of<T>(*a.t), a.coeffs, (T*)a.spec, a.spec_stride * 2, a.coeff_stride,
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
