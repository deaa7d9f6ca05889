// Colatitude (Legendre) and radial contractions as FP64 MFMA GEMMs.
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
// the radial stages. A workgroup of four waves computes a 32x64 tile of one
// problem with v_mfma_f64_16x16x4_f64; operand tiles are staged through LDS
// in chunks of 32 along the contraction axis. Work lists of (problem, row
// tile) are built per order on the host; grid.y walks the column tiles and
// grid.z the instances of the chunk.
#include <hip/hip_runtime.h>

#include "zgl_internal.hpp"

namespace zgl {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int TM = 32, TN = 64, TK = 32;

enum LoadOrder { K_FAST = 0, MN_FAST = 1, PAIR_K_FAST = 2 };

// One 32x64 output tile; S supplies the problem geometry and operand maps.
template <class S>
__global__ __launch_bounds__(256) void gemm_kernel(S st, const int4* __restrict__ work) {
    const int4 wk = work[blockIdx.x];
    const int n0 = blockIdx.y * TN;
    const typename S::Prob P = st.problem(wk.x, blockIdx.z);
    const int m0 = wk.y;
    if (!P.active || m0 >= P.M || n0 >= P.N) return;

    __shared__ double As[TM][TK + 1];
    __shared__ double Bs[TK][TN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    dbl4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};

    for (int k0 = 0; k0 < P.K; k0 += TK) {
        for (int e = tid; e < TM * TK; e += 256) {
            int i, k;
            if (S::A_ORDER == K_FAST) { k = e % TK; i = e / TK; }
            else { i = e % TM; k = e / TM; }
            const int gi = m0 + i, gk = k0 + k;
            As[i][k] = (gi < P.M && gk < P.K) ? st.a(P, gi, gk) : 0.0;
        }
        for (int e = tid; e < TK * TN; e += 256) {
            int k, j;
            if (S::B_ORDER == MN_FAST) { j = e % TN; k = e / TN; }
            else if (S::B_ORDER == K_FAST) { k = e % TK; j = e / TK; }
            else { const int ri = e & 1; k = (e >> 1) % TK; j = 2 * ((e >> 1) / TK) + ri; }
            const int gk = k0 + k, gj = n0 + j;
            Bs[k][j] = (gk < P.K && gj < P.N) ? st.b(P, gk, gj) : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; kk += 4) {
            const int kr = kk + (lane >> 4);
            const double bv = Bs[kr][wave * 16 + (lane & 15)];
            const double a0 = As[lane & 15][kr];
            const double a1 = As[16 + (lane & 15)][kr];
            acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bv, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bv, acc1, 0, 0, 0);
        }
        __syncthreads();
    }
    const int j = n0 + wave * 16 + (lane & 15);
    if (j >= P.N) return;
    // v_mfma_f64_16x16x4_f64 result: lane holds D[lane/16 + 4r][lane%16], r < 4
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i0 = m0 + (lane >> 4) + 4 * r;
        if (i0 < P.M) st.c(P, i0, j, acc0[r]);
        const int i1 = i0 + 16;
        if (i1 < P.M) st.c(P, i1, j, acc1[r]);
    }
}

// ---------------------------------------------------------------- Zernike
// Legendre forward: flm[l,m][k] = sum_h P_lm(z_h) X^par_m[h][k],
// l = m + par + 2i (rows i), columns (k, re/im).
struct ZtLegendreFwd {
    static constexpr int A_ORDER = K_FAST, B_ORDER = MN_FAST;
    const double* ptab; const double* spec; double* flm;
    long spec_stride, flm_stride;  // doubles
    int order, nt, jh, ncol;
    struct Prob { int active, M, N, K, m, par; const double* A; const double* B; double* C; };
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
    __device__ double a(const Prob& p, int i, int k) const { return p.A[(long)i * jh + k]; }
    __device__ double b(const Prob& p, int k, int j) const { return p.B[(long)k * 2 * ncol + j]; }
    __device__ void c(const Prob& p, int i, int j, double v) const {
        const int l = p.m + p.par + 2 * i;
        p.C[tri_idx(l, p.m) * 2 * ncol + j] = v;
    }
};

// Legendre backward: X^par_m[h][k] = sum_i P_lm(z_h) flm[l,m][k].
struct ZtLegendreBwd {
    static constexpr int A_ORDER = MN_FAST, B_ORDER = MN_FAST;
    const double* ptab; const double* flm; double* spec;
    long spec_stride, flm_stride;
    int order, nt, jh, ncol;
    struct Prob { int active, M, N, K, m, par; const double* A; const double* B; double* C; };
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
    __device__ double a(const Prob& p, int i, int k) const { return p.A[(long)k * jh + i]; }
    __device__ double b(const Prob& p, int k, int j) const {
        const int l = p.m + p.par + 2 * k;
        return p.B[tri_idx(l, p.m) * 2 * ncol + j];
    }
    __device__ void c(const Prob& p, int i, int j, double v) const { p.C[(long)i * 2 * ncol + j] = v; }
};

// Radial forward: c[n,l,m] = kappa_n sum_k R_nl(r_k) flm[l,m][k], n = l + 2i.
struct ZtRadialFwd {
    static constexpr int A_ORDER = K_FAST, B_ORDER = PAIR_K_FAST;
    const double* rtab; const double* flm; double* coeffs;
    long flm_stride, coeff_stride;
    int nt, nr, unnormed;
    struct Prob { int active, M, N, K, l; const double* B; double* C; };
    __device__ Prob problem(int prob, int inst) const {
        Prob p;
        p.l = prob;
        p.M = (nt - p.l + 1) / 2; p.K = nr; p.N = 2 * (p.l + 1);
        p.active = p.l < nt;
        p.B = flm + inst * flm_stride + tri_idx(p.l, 0) * 2 * nr;
        p.C = coeffs + inst * coeff_stride;
        return p;
    }
    __device__ double a(const Prob& p, int i, int k) const {
        return rtab[rad_idx(p.l + 2 * i, p.l) * nr + k];
    }
    __device__ double b(const Prob& p, int k, int j) const {
        return p.B[(long)(j >> 1) * 2 * nr + 2 * k + (j & 1)];
    }
    __device__ void c(const Prob& p, int i, int j, double v) const {
        const int n = p.l + 2 * i;
        const double kappa = unnormed ? (double)(2 * n + 3) : 1.0;
        p.C[tet_idx(n, p.l, 0) * 2 + j] = v * kappa;
    }
};

// Radial backward: flm[l,m][k] = sum_i R_nl(r_k) c[n,l,m], n = l + 2i;
// rows (m, re/im), columns k.
struct ZtRadialBwd {
    static constexpr int A_ORDER = MN_FAST, B_ORDER = MN_FAST;
    const double* rtab; const double* coeffs; double* flm;
    long flm_stride, coeff_stride;
    int nt, nr;
    struct Prob { int active, M, N, K, l; const double* A; double* C; };
    __device__ Prob problem(int prob, int inst) const {
        Prob p;
        p.l = prob;
        p.M = 2 * (p.l + 1); p.K = (nt - p.l + 1) / 2; p.N = nr;
        p.active = p.l < nt;
        p.A = coeffs + inst * coeff_stride;
        p.C = flm + inst * flm_stride + tri_idx(p.l, 0) * 2 * nr;
        return p;
    }
    __device__ double a(const Prob& p, int i, int k) const {
        return p.A[tet_idx(p.l + 2 * k, p.l, 0) * 2 + i];
    }
    __device__ double b(const Prob& p, int k, int j) const {
        return rtab[rad_idx(p.l + 2 * k, p.l) * nr + j];
    }
    __device__ void c(const Prob& p, int i, int j, double v) const {
        p.C[(long)(i >> 1) * 2 * nr + 2 * j + (i & 1)] = v;
    }
};

// ------------------------------------------------------ spherical harmonics
// The instances of a chunk form the column axis: columns (instance, re/im).
struct StLegendreFwd {
    static constexpr int A_ORDER = K_FAST, B_ORDER = PAIR_K_FAST;
    const double* ptab; const double* spec; double* coeffs;
    long spec_stride, coeff_stride;  // doubles
    int order, nt, jh, nb;
    struct Prob { int active, M, N, K, m, par; const double* A; long boff; };
    __device__ Prob problem(int prob, int) const {
        Prob p;
        p.m = prob >> 1; p.par = prob & 1;
        p.M = plm_count(nt, p.m, p.par); p.K = jh; p.N = 2 * nb;
        p.active = p.M > 0;
        p.A = ptab + plm_offset(order, p.m, p.par, jh);
        p.boff = (long)(p.m * 2 + p.par) * jh * 2;
        return p;
    }
    __device__ double a(const Prob& p, int i, int k) const { return p.A[(long)i * jh + k]; }
    __device__ double b(const Prob& p, int k, int j) const {
        return spec[(long)(j >> 1) * spec_stride + p.boff + 2 * k + (j & 1)];
    }
    __device__ void c(const Prob& p, int i, int j, double v) const {
        const int l = p.m + p.par + 2 * i;
        coeffs[(long)(j >> 1) * coeff_stride + tri_idx(l, p.m) * 2 + (j & 1)] = v;
    }
};

struct StLegendreBwd {
    static constexpr int A_ORDER = MN_FAST, B_ORDER = MN_FAST;
    const double* ptab; const double* coeffs; double* spec;
    long spec_stride, coeff_stride;
    int order, nt, jh, nb;
    struct Prob { int active, M, N, K, m, par; const double* A; long soff; };
    __device__ Prob problem(int prob, int) const {
        Prob p;
        p.m = prob >> 1; p.par = prob & 1;
        p.M = jh; p.K = plm_count(nt, p.m, p.par); p.N = 2 * nb;
        p.active = p.m < nt;
        p.A = ptab + plm_offset(order, p.m, p.par, jh);
        p.soff = (long)(p.m * 2 + p.par) * jh * 2;
        return p;
    }
    __device__ double a(const Prob& p, int i, int k) const { return p.A[(long)k * jh + i]; }
    __device__ double b(const Prob& p, int k, int j) const {
        const int l = p.m + p.par + 2 * k;
        return coeffs[(long)(j >> 1) * coeff_stride + tri_idx(l, p.m) * 2 + (j & 1)];
    }
    __device__ void c(const Prob& p, int i, int j, double v) const {
        spec[(long)(j >> 1) * spec_stride + p.soff + 2 * i + (j & 1)] = v;
    }
};

template <class S>
void launch(const S& st, const int4* work, int nwork, int ncols, int nz, hipStream_t s) {
    if (nwork == 0 || ncols <= 0 || nz <= 0) return;
    dim3 grid(nwork, (ncols + TN - 1) / TN, nz);
    gemm_kernel<S><<<grid, 256, 0, s>>>(st, work);
    check_hip(hipGetLastError(), "gemm launch");
}

}  // namespace

void zt_legendre_forward(const ZtGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    ZtLegendreFwd st{a.t->ptab, (const double*)a.spec, (double*)a.flm, a.spec_stride * 2,
                     a.flm_stride * 2, g.order, a.nt, g.jh, g.ncol};
    launch(st, a.t->work_lf, a.t->n_lf, 2 * g.ncol, a.nb, s);
}

void zt_legendre_backward(const ZtGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    ZtLegendreBwd st{a.t->ptab, (const double*)a.flm, (double*)a.spec, a.spec_stride * 2,
                     a.flm_stride * 2, g.order, a.nt, g.jh, g.ncol};
    launch(st, a.t->work_lb, a.t->n_lb, 2 * g.ncol, a.nb, s);
}

void zt_radial_forward(const ZtGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    ZtRadialFwd st{a.t->rtab, (const double*)a.flm, a.coeffs, a.flm_stride * 2, a.coeff_stride,
                   a.nt, g.ncol, a.zernike_unnormed};
    launch(st, a.t->work_rf, a.t->n_rf, 2 * g.order, a.nb, s);
}

void zt_radial_backward(const ZtGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    ZtRadialBwd st{a.t->rtab, a.coeffs, (double*)a.flm, a.flm_stride * 2, a.coeff_stride, a.nt,
                   g.ncol};
    launch(st, a.t->work_rb, a.t->n_rb, g.ncol, a.nb, s);
}

void st_legendre_forward(const StGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    StLegendreFwd st{a.t->ptab, (const double*)a.spec, a.coeffs, a.spec_stride * 2,
                     a.coeff_stride, g.order, a.nt, g.jh, a.nb};
    launch(st, a.t->work_lf, a.t->n_lf, 2 * a.nb, 1, s);
}

void st_legendre_backward(const StGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    StLegendreBwd st{a.t->ptab, a.coeffs, (double*)a.spec, a.spec_stride * 2, a.coeff_stride,
                     g.order, a.nt, g.jh, a.nb};
    launch(st, a.t->work_lb, a.t->n_lb, 2 * a.nb, 1, s);
}

}  // namespace zgl
