// Longitudinal pass: batched real DFTs along the longitude axis, for any
// length (odd lengths 2N-1 are often prime), fused with the quadrature
// weights, the conjugation and the north/south parity split (forward), or
// with the parity merge, the m > 0 halving and the Hermitian extension
// (backward).
//
// Replaces pocketfft::r2c / c2r at the call sites
//   include/zernike_glq_transformer.hpp:710-776 (r2c + apply_weights)
//   include/zernike_glq_transformer.hpp:943-949 (c2r)
//   include/sh_glq_transformer.hpp:757-884, 1136-1201 (r2c + fft_to_symm_asymm,
//                                                     symm_asymm_to_fft + c2r)
//
// Design. A workgroup owns T "column pairs": the columns of a latitude j
// and of its mirror nth-1-j at the same radial node. The two real columns are
// packed as one complex sequence z = a + i b (the classic two-for-one real
// FFT), so one complex FFT yields both spectra, and the pair is exactly what
// the parity split needs. The T sequences live in LDS as [n][T] complex
// doubles (lanes of a wave touch consecutive columns: conflict-free), and are
// transformed in place by a mixed-radix decimation-in-time FFT whose input is
// stored in digit-reversed order during the load. Each pass of radix p runs
// p-point DFTs in registers (the symmetric cos/sin form for odd p) and folds
// the next pass's twiddle factors into its stores. Prime radices above 17 use
// an LDS-staged variant that computes two outputs per work item.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "zgl_internal.hpp"

namespace zgl {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBigItems = 9;
constexpr int kMaxRegRadix = 17;  // larger primes use the LDS-staged pass

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Twiddle of the next pass for LDS slot `pos`: in a pass of radix pn whose
// sub-transforms have length Ln, slot pos = B*pn*Ln + r*Ln + q takes
// w^(r q) with w = e^{-/+ 2 pi i / (pn Ln)}.
template <int DIR>
__device__ inline double2 next_twiddle(const double2* twl, int n, int pos, int Ln, int pn) {
    const int q = pos % Ln;
    const int r = (pos / Ln) % pn;
    const int idx = (r * q * (n / (Ln * pn))) % n;
    double2 w = twl[idx];
    if (DIR > 0) w.y = -w.y;
    return w;
}

// In-register p-point DFT, y_s = sum_r x_r e^{DIR 2 pi i r s / p}.
template <int P, int DIR>
__device__ inline void dft_small(double2 (&x)[P], const double2* twl, int n) {
    if constexpr (P == 2) {
        const double2 a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    } else if constexpr (P == 4) {
        const double2 a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
        const double2 c = cadd(x[1], x[3]), d = csub(x[1], x[3]);
        // forward: y1 = b - i d, y3 = b + i d
        const double2 mid = make_double2(d.y, -d.x);  // -i d
        x[0] = cadd(a, c);
        x[2] = csub(a, c);
        if (DIR < 0) {
            x[1] = cadd(b, mid);
            x[3] = csub(b, mid);
        } else {
            x[1] = csub(b, mid);
            x[3] = cadd(b, mid);
        }
    } else {
        constexpr int H = (P - 1) / 2;
        double cs[H + 1], sn[H + 1];
#pragma unroll
        for (int t = 1; t <= H; ++t) {
            const double2 w = twl[t * (n / P)];  // e^{-2 pi i t / P}
            cs[t] = w.x;
            sn[t] = -w.y;
        }
        const double2 x0 = x[0];
        double2 y0 = x0;
#pragma unroll
        for (int r = 1; r <= H; ++r) {
            const double2 u = cadd(x[r], x[P - r]);
            const double2 v = csub(x[r], x[P - r]);
            x[r] = u;
            x[P - r] = v;
            y0 = cadd(y0, u);
        }
        double2 ys[H + 1], yc[H + 1];
#pragma unroll
        for (int s = 1; s <= H; ++s) {
            double2 a = x0, b = make_double2(0.0, 0.0);
#pragma unroll
            for (int r = 1; r <= H; ++r) {
                const int t = (r * s) % P;
                const double c = (t <= H) ? cs[t] : cs[P - t];
                const double si = (t <= H) ? sn[t] : -sn[P - t];
                a.x += x[r].x * c;
                a.y += x[r].y * c;
                b.x += x[P - r].x * si;
                b.y += x[P - r].y * si;
            }
            // forward: y_s = a - i b, y_{P-s} = a + i b
            const double2 mib = make_double2(b.y, -b.x);
            if (DIR < 0) {
                ys[s] = cadd(a, mib);
                yc[s] = csub(a, mib);
            } else {
                ys[s] = csub(a, mib);
                yc[s] = cadd(a, mib);
            }
        }
        x[0] = y0;
#pragma unroll
        for (int s = 1; s <= H; ++s) {
            x[s] = ys[s];
            x[P - s] = yc[s];
        }
    }
}

template <int P, int T, int DIR>
__device__ void pass_reg(double2* data, const double2* twl, int n, int L, int pn, int Ln) {
    const int nbf = n / P;
    for (int it = threadIdx.x; it < nbf * T; it += kThreads) {
        const int t = it % T;
        const int bf = it / T;
        const int bq = bf / L;
        const int q = bf - bq * L;
        const int base = bq * P * L + q;
        double2 x[P];
#pragma unroll
        for (int r = 0; r < P; ++r) x[r] = data[(base + r * L) * T + t];
        dft_small<P, DIR>(x, twl, n);
#pragma unroll
        for (int s = 0; s < P; ++s) {
            const int pos = base + s * L;
            double2 y = x[s];
            if (pn) y = cmul(y, next_twiddle<DIR>(twl, n, pos, Ln, pn));
            data[pos * T + t] = y;
        }
    }
}

// Prime radix P > 17: work item = (butterfly, transform, s in [0, H]) producing
// y_s and y_{P-s}; results are held in registers across a barrier because
// the DFT reads every input slot of its butterfly.
template <int T, int DIR>
__device__ void pass_big(double2* data, const double2* twl, int n, int P, int L, int pn, int Ln) {
    const int H = (P - 1) / 2;
    const int nbf = n / P;
    const int items = nbf * T * (H + 1);
    const int stride = n / P;
    double2 ya[kMaxBigItems], yb[kMaxBigItems];
#pragma unroll
    for (int c = 0; c < kMaxBigItems; ++c) {
        const int it = threadIdx.x + c * kThreads;
        if (it >= items) break;
        const int s = it % (H + 1);
        const int rest = it / (H + 1);
        const int t = rest % T;
        const int bf = rest / T;
        const int bq = bf / L;
        const int q = bf - bq * L;
        const int base = bq * P * L + q;
        const double2 x0 = data[base * T + t];
        if (s == 0) {
            double2 y = x0;
            for (int r = 1; r < P; ++r) y = cadd(y, data[(base + r * L) * T + t]);
            ya[c] = y;
            yb[c] = y;
        } else {
            double2 a = x0, b = make_double2(0.0, 0.0);
            int idx = 0;
            for (int r = 1; r <= H; ++r) {
                const double2 xr = data[(base + r * L) * T + t];
                const double2 xm = data[(base + (P - r) * L) * T + t];
                idx += s;
                if (idx >= P) idx -= P;
                const double2 w = twl[idx * stride];
                const double c0 = w.x, s0 = -w.y;
                a.x += (xr.x + xm.x) * c0;
                a.y += (xr.y + xm.y) * c0;
                b.x += (xr.x - xm.x) * s0;
                b.y += (xr.y - xm.y) * s0;
            }
            const double2 mib = make_double2(b.y, -b.x);
            if (DIR < 0) {
                ya[c] = cadd(a, mib);
                yb[c] = csub(a, mib);
            } else {
                ya[c] = csub(a, mib);
                yb[c] = cadd(a, mib);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kMaxBigItems; ++c) {
        const int it = threadIdx.x + c * kThreads;
        if (it >= items) break;
        const int s = it % (H + 1);
        const int rest = it / (H + 1);
        const int t = rest % T;
        const int bf = rest / T;
        const int bq = bf / L;
        const int q = bf - bq * L;
        const int base = bq * P * L + q;
        {
            const int pos = base + s * L;
            double2 y = ya[c];
            if (pn) y = cmul(y, next_twiddle<DIR>(twl, n, pos, Ln, pn));
            data[pos * T + t] = y;
        }
        if (s > 0) {
            const int pos = base + (P - s) * L;
            double2 y = yb[c];
            if (pn) y = cmul(y, next_twiddle<DIR>(twl, n, pos, Ln, pn));
            data[pos * T + t] = y;
        }
    }
}

template <int T, int DIR>
__device__ void run_passes(double2* data, const double2* twl, const FftDesc& d) {
    const int n = d.n;
    for (int ps = 0; ps < d.npass; ++ps) {
        __syncthreads();
        const int P = d.radix[ps], L = d.L[ps];
        const int pn = ps + 1 < d.npass ? d.radix[ps + 1] : 0;
        const int Ln = L * P;
        switch (P) {
            case 2: pass_reg<2, T, DIR>(data, twl, n, L, pn, Ln); break;
            case 3: pass_reg<3, T, DIR>(data, twl, n, L, pn, Ln); break;
            case 4: pass_reg<4, T, DIR>(data, twl, n, L, pn, Ln); break;
            case 5: pass_reg<5, T, DIR>(data, twl, n, L, pn, Ln); break;
            case 7: pass_reg<7, T, DIR>(data, twl, n, L, pn, Ln); break;
            case 11: pass_reg<11, T, DIR>(data, twl, n, L, pn, Ln); break;
            case 13: pass_reg<13, T, DIR>(data, twl, n, L, pn, Ln); break;
            case 17: pass_reg<17, T, DIR>(data, twl, n, L, pn, Ln); break;
            default: pass_big<T, DIR>(data, twl, n, P, L, pn, Ln); break;
        }
    }
    __syncthreads();
}

struct PairMap {
    int npairs, ncol, nth, jh, lat_stride, lon_stride;
};

template <int T, int LK>
__global__ __launch_bounds__(kThreads) void fft_forward_kernel(
    FftDesc d, const double2* __restrict__ twg, const unsigned short* __restrict__ permg,
    const double* __restrict__ grid, long grid_stride, double2* __restrict__ spec,
    long spec_stride, PairMap pm, int nm, double scale, const double* __restrict__ wlat,
    const double* __restrict__ wrad, double center_factor) {
    extern __shared__ double2 smem[];
    const int n = d.n;
    double2* data = smem;
    double2* twl = smem + n * T;
    const int tile = blockIdx.x;
    const double* g = grid + (long)blockIdx.y * grid_stride;
    double2* sp = spec + (long)blockIdx.y * spec_stride;
    for (int i = threadIdx.x; i < n; i += kThreads) twl[i] = twg[i];

    if (LK == GRID_COLS) {
        for (int e = threadIdx.x; e < n * T; e += kThreads) {
            const int i = e / T, t = e % T;
            const int p = tile * T + t;
            double a = 0.0, b = 0.0;
            if (p < pm.npairs) {
                const int h = p / pm.ncol, c = p - h * pm.ncol;
                const int jn = pm.nth / 2 + h, js = pm.jh - 1 - h;
                const double* row = g + (long)i * pm.lon_stride + c;
                a = row[(long)jn * pm.lat_stride];
                if (js != jn) b = row[(long)js * pm.lat_stride];
            }
            data[permg[i] * T + t] = make_double2(a, b);
        }
    } else {
        for (int e = threadIdx.x; e < n * T; e += kThreads) {
            const int t = e / n, i = e - t * n;
            const int p = tile * T + t;
            double a = 0.0, b = 0.0;
            if (p < pm.npairs) {
                const int jn = pm.nth / 2 + p, js = pm.jh - 1 - p;
                a = g[(long)jn * pm.lat_stride + i];
                if (js != jn) b = g[(long)js * pm.lat_stride + i];
            }
            data[permg[i] * T + t] = make_double2(a, b);
        }
    }

    run_passes<T, -1>(data, twl, d);

    for (int e = threadIdx.x; e < nm * T; e += kThreads) {
        const int m = e / T, t = e % T;
        const int p = tile * T + t;
        if (p >= pm.npairs) continue;
        const double2 zm = data[m * T + t];
        const double2 zc = data[((n - m) % n) * T + t];
        // two-for-one split: A = (Z_m + conj Z_{n-m}) / 2, B = (Z_m - conj Z_{n-m}) / (2i)
        const double2 A = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y - zc.y));
        const double2 B = make_double2(0.5 * (zm.y + zc.y), -0.5 * (zm.x - zc.x));
        const int h = p / pm.ncol, c = p - h * pm.ncol;
        const double W = wlat[h] * (wrad ? wrad[c] : 1.0) * scale;
        // H = weight * conj(scale * X)
        const double2 Ha = make_double2(W * A.x, -W * A.y);
        const double2 Hb = make_double2(W * B.x, -W * B.y);
        double2 E, O;
        const bool center = (pm.nth / 2 + h) == (pm.jh - 1 - h);
        if (center) {
            E = make_double2(center_factor * Ha.x, center_factor * Ha.y);
            O = make_double2(0.0, 0.0);
        } else {
            E = cadd(Ha, Hb);
            O = csub(Ha, Hb);
        }
        sp[((long)(m * 2 + 0) * pm.jh + h) * pm.ncol + c] = E;
        sp[((long)(m * 2 + 1) * pm.jh + h) * pm.ncol + c] = O;
    }
}

template <int T, int LK>
__global__ __launch_bounds__(kThreads) void fft_backward_kernel(
    FftDesc d, const double2* __restrict__ twg, const unsigned short* __restrict__ permg,
    const double2* __restrict__ spec, long spec_stride, double* __restrict__ grid,
    long grid_stride, PairMap pm, int nm) {
    extern __shared__ double2 smem[];
    const int n = d.n;
    double2* data = smem;
    double2* twl = smem + n * T;
    const int tile = blockIdx.x;
    double* g = grid + (long)blockIdx.y * grid_stride;
    const double2* sp = spec + (long)blockIdx.y * spec_stride;
    for (int i = threadIdx.x; i < n; i += kThreads) twl[i] = twg[i];
    const int F = n / 2 + 1;
    for (int e = threadIdx.x; e < F * T; e += kThreads) {
        const int m = e / T, t = e % T;
        const int p = tile * T + t;
        double2 zm = make_double2(0.0, 0.0), zp = make_double2(0.0, 0.0);
        if (p < pm.npairs && m < nm) {
            const int h = p / pm.ncol, c = p - h * pm.ncol;
            const double2 Ae = sp[((long)(m * 2 + 0) * pm.jh + h) * pm.ncol + c];
            const double2 Ao = sp[((long)(m * 2 + 1) * pm.jh + h) * pm.ncol + c];
            const double mu = m > 0 ? 0.5 : 1.0;
            const bool center = (pm.nth / 2 + h) == (pm.jh - 1 - h);
            double2 Sa, Sb;
            if (center) {
                Sa = make_double2(mu * Ae.x, -mu * Ae.y);
                Sb = make_double2(0.0, 0.0);
            } else {
                const double2 Ev = cadd(Ae, Ao), Dv = csub(Ae, Ao);
                Sa = make_double2(mu * Ev.x, -mu * Ev.y);
                Sb = make_double2(mu * Dv.x, -mu * Dv.y);
            }
            if (m == 0 || 2 * m == n) {
                zm = make_double2(Sa.x, Sb.x);  // imaginary parts of X_0 (and Nyquist) ignored
            } else {
                zm = make_double2(Sa.x - Sb.y, Sa.y + Sb.x);   // Sa + i Sb
                zp = make_double2(Sa.x + Sb.y, -Sa.y + Sb.x);  // conj(Sa) + i conj(Sb)
            }
        }
        data[permg[m] * T + t] = zm;
        if (m > 0 && 2 * m != n) data[permg[n - m] * T + t] = zp;
    }

    run_passes<T, +1>(data, twl, d);

    if (LK == GRID_COLS) {
        for (int e = threadIdx.x; e < n * T; e += kThreads) {
            const int i = e / T, t = e % T;
            const int p = tile * T + t;
            if (p >= pm.npairs) continue;
            const double2 z = data[i * T + t];
            const int h = p / pm.ncol, c = p - h * pm.ncol;
            const int jn = pm.nth / 2 + h, js = pm.jh - 1 - h;
            double* row = g + (long)i * pm.lon_stride + c;
            row[(long)jn * pm.lat_stride] = z.x;
            if (js != jn) row[(long)js * pm.lat_stride] = z.y;
        }
    } else {
        for (int e = threadIdx.x; e < n * T; e += kThreads) {
            const int t = e / n, i = e - t * n;
            const int p = tile * T + t;
            if (p >= pm.npairs) continue;
            const double2 z = data[i * T + t];
            const int jn = pm.nth / 2 + p, js = pm.jh - 1 - p;
            g[(long)jn * pm.lat_stride + i] = z.x;
            if (js != jn) g[(long)js * pm.lat_stride + i] = z.y;
        }
    }
}

int lds_limit() {
    static int limit = 0;
    if (!limit) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
            v > 0)
            limit = v;
        else
            limit = 65536;
    }
    return limit;
}

size_t lds_bytes(int n, int T) { return (size_t)(n * T + n) * sizeof(double2); }

template <int T>
void launch_forward(const FftLaunch& L, const PairMap& pm, const double* grid, long gs,
                    double2* spec, long ss, int nb, hipStream_t s) {
    const FftPlan& f = L.t->fft;
    const int ntiles = (pm.npairs + T - 1) / T;
    const size_t lds = lds_bytes(f.desc.n, T);
    dim3 grd(ntiles, nb);
    if (L.layout_kind == GRID_COLS) {
        auto k = fft_forward_kernel<T, GRID_COLS>;
        if (lds > 65536) check_hip(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds), "lds attribute");
        k<<<grd, kThreads, lds, s>>>(f.desc, f.tw, f.perm, grid, gs, spec, ss, pm, L.nm_used,
                                     L.scale, L.t->wlat, L.t->wrad, L.center_factor);
    } else {
        auto k = fft_forward_kernel<T, GRID_ROWS>;
        if (lds > 65536) check_hip(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds), "lds attribute");
        k<<<grd, kThreads, lds, s>>>(f.desc, f.tw, f.perm, grid, gs, spec, ss, pm, L.nm_used,
                                     L.scale, L.t->wlat, L.t->wrad, L.center_factor);
    }
    check_hip(hipGetLastError(), "fft_forward launch");
}

template <int T>
void launch_backward(const FftLaunch& L, const PairMap& pm, const double2* spec, long ss,
                     double* grid, long gs, int nb, hipStream_t s) {
    const FftPlan& f = L.t->fft;
    const int ntiles = (pm.npairs + T - 1) / T;
    const size_t lds = lds_bytes(f.desc.n, T);
    dim3 grd(ntiles, nb);
    if (L.layout_kind == GRID_COLS) {
        auto k = fft_backward_kernel<T, GRID_COLS>;
        if (lds > 65536) check_hip(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds), "lds attribute");
        k<<<grd, kThreads, lds, s>>>(f.desc, f.tw, f.perm, spec, ss, grid, gs, pm, L.nm_used);
    } else {
        auto k = fft_backward_kernel<T, GRID_ROWS>;
        if (lds > 65536) check_hip(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds), "lds attribute");
        k<<<grd, kThreads, lds, s>>>(f.desc, f.tw, f.perm, spec, ss, grid, gs, pm, L.nm_used);
    }
    check_hip(hipGetLastError(), "fft_backward launch");
}

PairMap pair_map(const Geom& g) {
    PairMap pm;
    pm.ncol = g.ncol;
    pm.nth = g.nth;
    pm.jh = g.jh;
    pm.npairs = g.jh * g.ncol;
    pm.lat_stride = g.lat_stride;
    pm.lon_stride = g.lon_stride;
    return pm;
}

}  // namespace

void fft_plan_create(FftPlan& p, int n, hipStream_t s) {
    p = FftPlan{};
    if (n <= 0) return;
    if (n > 65535) throw Error{ZGL_ERROR_NOT_SUPPORTED, "longitude count above 65535"};
    // Radices: 4s, then 2, then odd primes; executed largest first.
    std::vector<int> rad;
    int m = n;
    while (m % 4 == 0) { rad.push_back(4); m /= 4; }
    if (m % 2 == 0) { rad.push_back(2); m /= 2; }
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0) { rad.push_back(f); m /= f; }
        if ((long)f * f > m && m > 1) { rad.push_back(m); m = 1; }
    }
    std::sort(rad.begin(), rad.end(), [](int a, int b) { return a > b; });
    if ((int)rad.size() > kMaxPass) throw Error{ZGL_ERROR_NOT_SUPPORTED, "too many FFT passes"};
    p.desc.n = n;
    p.desc.npass = (int)rad.size();
    int L = 1;
    for (size_t i = 0; i < rad.size(); ++i) {
        p.desc.radix[i] = rad[i];
        p.desc.L[i] = L;
        L *= rad[i];
    }
    // Tile width: the most column pairs whose LDS fits, and whose large-prime
    // passes fit the per-thread register budget.
    const int limit = std::min(lds_limit(), 160 * 1024);
    int T = 16;
    for (; T > 1; T /= 2) {
        if (lds_bytes(n, T) > (size_t)limit) continue;
        bool ok = true;
        for (int r : rad)
            if (r > kMaxRegRadix) {
                const long items = (long)(n / r) * T * ((r - 1) / 2 + 1);
                if ((items + kThreads - 1) / kThreads > kMaxBigItems) ok = false;
            }
        if (ok) break;
    }
    if (lds_bytes(n, T) > (size_t)limit)
        throw Error{ZGL_ERROR_NOT_SUPPORTED, "longitude count too large for the LDS"};
    p.tile = T;

    // Input slot of sample i for the DIT order: the last pass combines
    // sub-transforms of samples i = r (mod p_last), stored in block r.
    std::vector<unsigned short> perm(n);
    for (int i = 0; i < n; ++i) {
        int pos = 0, len = n, idx = i;
        for (int k = (int)rad.size() - 1; k >= 0; --k) {
            const int r = rad[k];
            len /= r;
            pos += (idx % r) * len;
            idx /= r;
        }
        perm[i] = (unsigned short)pos;
    }
    std::vector<double2> tw(n);
    for (int t = 0; t < n; ++t) {
        const long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)n;
        tw[t] = make_double2((double)cosl(ang), (double)(-sinl(ang)));
    }
    check_hip(hipMalloc(&p.tw, n * sizeof(double2)), "fft twiddles");
    check_hip(hipMalloc(&p.perm, n * sizeof(unsigned short)), "fft permutation");
    check_hip(hipMemcpyAsync(p.tw, tw.data(), n * sizeof(double2), hipMemcpyHostToDevice, s), "upload twiddles");
    check_hip(hipMemcpyAsync(p.perm, perm.data(), n * sizeof(unsigned short), hipMemcpyHostToDevice, s), "upload permutation");
    check_hip(hipStreamSynchronize(s), "fft plan");
}

void fft_plan_destroy(FftPlan& p) {
    if (p.tw) (void)hipFree(p.tw);
    if (p.perm) (void)hipFree(p.perm);
    p = FftPlan{};
}

void fft_forward(const FftLaunch& L, const double* grid, long grid_stride, double2* spec,
                 long spec_stride, int nb, hipStream_t s) {
    const PairMap pm = pair_map(*L.g);
    switch (L.t->fft.tile) {
        case 16: launch_forward<16>(L, pm, grid, grid_stride, spec, spec_stride, nb, s); break;
        case 8: launch_forward<8>(L, pm, grid, grid_stride, spec, spec_stride, nb, s); break;
        case 4: launch_forward<4>(L, pm, grid, grid_stride, spec, spec_stride, nb, s); break;
        case 2: launch_forward<2>(L, pm, grid, grid_stride, spec, spec_stride, nb, s); break;
        default: launch_forward<1>(L, pm, grid, grid_stride, spec, spec_stride, nb, s); break;
    }
}

void fft_backward(const FftLaunch& L, const double2* spec, long spec_stride, double* grid,
                  long grid_stride, int nb, hipStream_t s) {
    const PairMap pm = pair_map(*L.g);
    switch (L.t->fft.tile) {
        case 16: launch_backward<16>(L, pm, spec, spec_stride, grid, grid_stride, nb, s); break;
        case 8: launch_backward<8>(L, pm, spec, spec_stride, grid, grid_stride, nb, s); break;
        case 4: launch_backward<4>(L, pm, spec, spec_stride, grid, grid_stride, nb, s); break;
        case 2: launch_backward<2>(L, pm, spec, spec_stride, grid, grid_stride, nb, s); break;
        default: launch_backward<1>(L, pm, spec, spec_stride, grid, grid_stride, nb, s); break;
    }
}

}  // namespace zgl
