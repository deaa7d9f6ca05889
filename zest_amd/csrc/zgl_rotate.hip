// Rotations of Zernike and spherical-harmonic coefficients: Rotor::rotate and
// Rotor::polar_rotate (include/rotor.hpp:124-380) with the Wigner d-matrices
// at pi/2 of WignerdPiHalfCollection (src/wignerd_collection.cpp:29-104).
//
// A rotation acts on each degree-l block of (l+1) complex values v_m (the
// {cos, sin} pairs with m > 0 scaled by (1/sqrt2, -1/sqrt2)) as the product
//   e^{-i m alpha'} X(-90) e^{-i m beta'} X(+90) e^{-i m gamma'}
// (rotate_l, include/wignerd_collection.hpp:134-296), where X(+90) folds the
// sum over negative m' by the symmetry of a real function:
//   Re y_m = sum_{m'} c_{m'} d_{m,m'} Re v_{m'}   over (m + m' + l) even,
//   Im y_m = sum_{m'} 2 d_{m,m'} Im v_{m'}        over (m + m' + l) odd,
// c_0 = 1, c_{m'>0} = 2, and X(-90) the same with d_{m',m}. Splitting the
// rows by the parity of m makes each of the four masked products a dense
// GEMM of about (l/2) x (l/2) against the columns: all blocks of degree l of
// the batch (SH: one per instance; Zernike: one per instance and n = l, l+2,
// ...). Two kernels, one per X step, each an FP64-MFMA GEMM with the phases
// fused into the operand staging and the store.
#include <hip/hip_runtime.h>

#include <cmath>

#include "zgl_internal.hpp"

namespace zgl {

namespace {

typedef double d2 __attribute__((ext_vector_type(2)));
typedef double d4 __attribute__((ext_vector_type(4)));

constexpr double kInvSqrt2 = 0.70710678118654752440;  // 1/sqrt2 rounded
constexpr double kSqrt2 = 1.41421356237309504880;

__device__ __host__ inline long wig_base(long l) { return l * (l + 1) * (2 * l + 1) / 6; }

// ---------------------------------------------------------------- d table
// One thread per (l, m1): the column d^l_{m2,m1}, m2 = m1..l, by the
// recursion downwards in m2 from d^l_{l,m1}, whose seed is formed by the same
// products as the reference's running value; the transposed entries by the
// symmetry d_{m1,m2} = (-1)^{m1+m2} d_{m2,m1}. Column 0 steps by 2 (its
// entries with l + m2 odd are zero); the table is zeroed beforehand.
__global__ void wignerd_kernel(int order, double* d) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    // t = l(l+1)/2 + m1
    int l = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((long)(l + 1) * (l + 2) / 2 <= t) ++l;
    while ((long)l * (l + 1) / 2 > t) --l;
    const int m1 = (int)(t - (long)l * (l + 1) / 2);
    if (l >= order) return;
    double* dl = d + wig_base(l);
    auto D = [&](int a, int b) -> double& { return dl[(long)(l + 1) * a + b]; };
    auto sq = [](int i) { return sqrt((double)i); };
    auto isq = [](int i) { return 1.0 / sqrt((double)i); };
    if (l == 0) {
        D(0, 0) = 1.0;
        return;
    }
    if (l == 1) {
        const double h = 1.0 / 1.41421356237309504880;
        if (m1 == 0) {
            D(0, 0) = 0.0;
            D(1, 0) = -h;
        } else {
            D(0, 1) = h;
            D(1, 1) = 0.5;
        }
        return;
    }
    double dl0 = -1.0 / 1.41421356237309504880;
    for (int k = 2; k <= l; ++k) dl0 *= -sq(2 * k - 1) / sq(2 * k);
    if (m1 == 0) {
        D(l, 0) = dl0;
        D(l - 1, 0) = 0.0;
        for (int i = 2; i <= l; i += 2) {
            const int m2 = l - i;
            D(m2, 0) = -D(m2 + 2, 0) * sq(l - m2 - 1) * sq(l + m2 + 2) * isq(l - m2) * isq(l + m2 + 1);
        }
        for (int m2 = 1; m2 <= l; ++m2) D(0, m2) = ((m2 & 1) ? -1.0 : 1.0) * D(m2, 0);
        return;
    }
    double dlm = dl0;
    const int last = m1 < l ? m1 : l - 1;
    for (int k = 1; k <= last; ++k) dlm *= -sq(l - k + 1) * isq(l + k);
    if (m1 == l) {
        D(l, l) = -dlm * isq(2 * l);
        return;
    }
    D(l, m1) = dlm;
    D(l - 1, m1) = D(l, m1) * 2.0 * (double)m1 * isq(2 * l);
    for (int i = 2; i <= l - m1; ++i) {
        const int m2 = l - i;
        D(m2, m1) = (2.0 * (double)m1 * D(m2 + 1, m1) - sq(l - m2 - 1) * sq(l + m2 + 2) * D(m2 + 2, m1)) *
                    isq(l - m2) * isq(l + m2 + 1);
    }
    for (int m2 = m1 + 1; m2 <= l; ++m2) D(m1, m2) = (((m1 ^ m2) & 1) ? -1.0 : 1.0) * D(m2, m1);
}

// ---------------------------------------------------------------- rotation
struct RotArgs {
    const double* dtab;
    const int4* work;          // (l, row parity, row tile, column tile)
    int zt, order, nb;
    Inst<const double> src;    // stage 0: the coefficients; stage 1: the intermediate (complex)
    Inst<double> dst;          // stage 0: the intermediate; stage 1: the coefficients
    const double* angles;      // [nb][3] alpha, beta, gamma as given
    int coordinate;            // RotationType::coordinate
    int stage;
};

// detail::convert (include/rotor.hpp:69-86): the angles the phases use
__device__ inline void converted(const RotArgs& a, int inst, double& al, double& be, double& ga) {
    const double* e = a.angles + 3L * inst;
    const double ar = e[0] - 0.5 * M_PI, br = e[1], gr = e[2] + 0.5 * M_PI;
    if (a.coordinate) {
        al = -gr; be = -br; ga = -ar;
    } else {
        al = ar; be = br; ga = gr;
    }
}

// v * e^{-i m x}, the reference's component order
__device__ inline d2 phase(d2 v, int m, double x) {
    double s, c;
    sincos(-(double)m * x, &s, &c);
    return d2{v[0] * c - v[1] * s, v[0] * s + v[1] * c};
}

constexpr int kRotRows = 64;   // rows of one parity per workgroup (16 per wave)
constexpr int kRotCols = 64;   // columns per workgroup (4 blocks of 16, every wave)
constexpr int kRotK = 64;      // m' per LDS chunk (32 of each parity)

__global__ __launch_bounds__(256) void rotate_kernel(RotArgs a) {
    __shared__ d2 Xs[kRotK * kRotCols];
    const int4 w = a.work[blockIdx.x];
    const int l = w.x, p = w.y;
    const int i0 = w.z * kRotRows, c0 = w.w * kRotCols;
    const int cnt = a.zt ? (a.order - l + 1) / 2 : 1;
    const int ncols = a.nb * cnt;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int j = lane & 15, g = lane >> 4;
    const int pa = (l + p) & 1, pb = pa ^ 1;  // m' parity of the Re and Im products
    const double* dl = a.dtab + wig_base(l);
    const int mrow = 2 * (i0 + 16 * wave + j) + p;  // A-operand row of this lane
    const bool rowok = mrow <= l;

    // pair index of (column, m')
    auto locate = [&](int col, int& inst) -> long {
        int q = 0;
        inst = col;
        if (a.zt) {
            inst = col / cnt;
            q = col - inst * cnt;
            return tet_idx(l + 2 * q, l, 0);
        }
        return tri_idx(l, 0);
    };

    d4 are[4], aim[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) are[c] = aim[c] = d4{0, 0, 0, 0};

    for (int k0 = 0; k0 <= l; k0 += kRotK) {
        __syncthreads();
        for (int e = tid; e < kRotK * kRotCols; e += 256) {
            const int kl = e / kRotCols, c = e - kl * kRotCols;
            const int col = c0 + c, mp = k0 + kl;
            d2 v = {0.0, 0.0};
            if (col < ncols && mp <= l) {
                int inst;
                const long base = locate(col, inst);
                if (a.stage == 0) {
                    const double* x = a.src.at(inst) + 2 * (base + mp);
                    v = mp ? d2{x[0] * kInvSqrt2, x[1] * -kInvSqrt2} : d2{x[0], x[1]};
                    double al, be, ga;
                    converted(a, inst, al, be, ga);
                    v = phase(v, mp, ga);
                } else {
                    v = *reinterpret_cast<const d2*>(a.src.at(inst) + 2 * (base + mp));
                }
            }
            Xs[e] = v;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kRotK / 8; ++s) {
            const int kk = 4 * s + g;
            const int mre = k0 + 2 * kk + pa, mim = k0 + 2 * kk + pb;
            double xa = 0.0, xb = 0.0;
            if (rowok && mre <= l)
                xa = (mre ? 2.0 : 1.0) * (a.stage == 0 ? dl[(long)(l + 1) * mrow + mre] : dl[(long)(l + 1) * mre + mrow]);
            if (rowok && mim <= l)
                xb = 2.0 * (a.stage == 0 ? dl[(long)(l + 1) * mrow + mim] : dl[(long)(l + 1) * mim + mrow]);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double bre = Xs[(2 * kk + pa) * kRotCols + 16 * c + j][0];
                const double bim = Xs[(2 * kk + pb) * kRotCols + 16 * c + j][1];
                are[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa, bre, are[c], 0, 0, 0);
                aim[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(xb, bim, aim[c], 0, 0, 0);
            }
        }
    }

#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int col = c0 + 16 * c + j;
        if (col >= ncols) continue;
        int inst;
        const long base = locate(col, inst);
        double al, be, ga;
        converted(a, inst, al, be, ga);
        double* out = a.dst.at(inst) + 2 * base;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = 2 * (i0 + 16 * wave + g + 4 * r) + p;  // D row of this lane
            if (m > l) continue;
            d2 y = {are[c][r], aim[c][r]};
            if (a.stage == 0) {
                y = phase(y, m, be);
            } else {
                y = phase(y, m, al);
                if (m) y = d2{y[0] * kSqrt2, y[1] * -kSqrt2};
            }
            *reinterpret_cast<d2*>(out + 2 * m) = y;
        }
    }
}

// ---------------------------------------------------------------- polar
// One thread per (instance, block): v_m <- e^{-i m angle'} v_m for m >= 1 in
// the complex form (Rotor::polar_rotate, include/rotor.hpp:321-380).
__global__ void polar_kernel(int zt, int order, long nblocks, Inst<double> c, const double* angles,
                             int coordinate) {
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= nblocks) return;
    const int inst = blockIdx.y;
    long first;
    int l;
    if (zt) {
        // blocks e = rad_idx(n, l), l >= 0
        int n = max(0, (int)(2.0 * sqrt((double)e)) - 2);
        while (((long)(n + 2) * (n + 2) >> 2) <= e) ++n;
        l = 2 * (int)(e - (((long)(n + 1) * (n + 1)) >> 2)) + (n & 1);
        first = tet_idx(n, l, 0);
    } else {
        l = (int)e;
        first = tri_idx(l, 0);
    }
    if (l == 0) return;
    const double x = coordinate ? -angles[inst] : angles[inst];
    double* v = c.at(inst) + 2 * first;
    for (int m = 1; m <= l; ++m) {
        d2 z = {v[2 * m] * kInvSqrt2, v[2 * m + 1] * -kInvSqrt2};
        z = phase(z, m, x);
        v[2 * m] = z[0] * kSqrt2;
        v[2 * m + 1] = z[1] * -kSqrt2;
    }
}

}  // namespace

long wignerd_size(int order) { return wig_base(order); }

void wignerd_table(int order, double* d, hipStream_t s) {
    const long n = wignerd_size(order);
    if (n == 0) return;
    check_hip(hipMemsetAsync(d, 0, n * sizeof(double), s), "d table clear");
    const long threads = (long)order * (order + 1) / 2;
    wignerd_kernel<<<(unsigned)((threads + 255) / 256), 256, 0, s>>>(order, d);
    check_hip(hipGetLastError(), "d table launch");
}

std::vector<int4> rotation_work(bool zt, int order, int nb) {
    std::vector<int4> w;
    for (int l = 1; l < order; ++l) {
        const int cnt = zt ? (order - l + 1) / 2 : 1;
        const long ncols = (long)nb * cnt;
        const int ct = (int)((ncols + kRotCols - 1) / kRotCols);
        for (int p = 0; p < 2; ++p) {
            if (p > l) continue;
            const int nrow = (l - p) / 2 + 1;
            const int rt = (nrow + kRotRows - 1) / kRotRows;
            for (int r = 0; r < rt; ++r)
                for (int c = 0; c < ct; ++c) w.push_back(make_int4(l, p, r, c));
        }
    }
    return w;
}

void rotate(bool zt, int order, const double* dtab, const int4* work, int nwork, Inst<double> coeffs,
            double* temp, long temp_stride, const double* angles, int coordinate, int nb, hipStream_t s) {
    if (nwork == 0 || nb == 0) return;
    RotArgs a{dtab, work, zt ? 1 : 0, order, nb, Inst<const double>(coeffs),
              Inst<double>{temp, temp_stride, nullptr}, angles, coordinate, 0};
    rotate_kernel<<<nwork, 256, 0, s>>>(a);
    check_hip(hipGetLastError(), "rotation launch");
    a.src = Inst<const double>{temp, temp_stride, nullptr};
    a.dst = coeffs;
    a.stage = 1;
    rotate_kernel<<<nwork, 256, 0, s>>>(a);
    check_hip(hipGetLastError(), "rotation launch");
}

void polar_rotate(bool zt, int order, Inst<double> coeffs, const double* angles, int coordinate, int nb,
                  hipStream_t s) {
    const long nblocks = zt ? (long)(order + 1) * (order + 1) / 4 : order;
    if (nblocks == 0 || nb == 0) return;
    polar_kernel<<<dim3((unsigned)((nblocks + 255) / 256), nb), 256, 0, s>>>(zt ? 1 : 0, order, nblocks, coeffs,
                                                                           angles, coordinate);
    check_hip(hipGetLastError(), "polar rotation launch");
}

}  // namespace zgl
