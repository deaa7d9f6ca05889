// Longitudinal pass: batched real DFTs along the longitude axis, for any
// length (odd lengths 2N-1 are often prime), fused with the quadrature
// weights, the conjugation and the north/south parity split (forward), or
// with the parity merge, the m > 0 halving and the Hermitian extension
// (backward). FP64 or FP32 arithmetic; FP64 grids, or FP32 grids with FP32
// arithmetic.
//
// Replaces pocketfft::r2c / c2r at the call sites
//   include/zernike_glq_transformer.hpp:710-776 (r2c + apply_weights)
//   include/zernike_glq_transformer.hpp:943-949 (c2r)
//   include/sh_glq_transformer.hpp:757-884, 1136-1201 (r2c + fft_to_symm_asymm,
//                                                     symm_asymm_to_fft + c2r)
//
// Design. A workgroup owns T "column pairs": the columns of a latitude j and
// of its mirror nth-1-j at the same radial node. The two real columns are
// packed as one complex sequence z = a + i b (the two-for-one real FFT), so
// one complex FFT yields both spectra, and the pair is exactly what the parity
// split needs. For Zernike grids a tile holds T consecutive entries of the
// flattened (latitude pair, radial node) index, so its loads are contiguous
// runs of up to T values per longitude row. The T sequences live in LDS as
// [n][T] complex values (lanes of a wave touch consecutive columns:
// conflict-free) and are transformed in place by a mixed-radix
// decimation-in-time FFT: the input is stored in digit-reversed order during
// the load, and each pass of radix p multiplies its inputs by a precomputed
// per-slot twiddle table (in LDS for wide tiles, read through the caches for
// narrow ones; no index arithmetic beyond the butterfly base) and runs p-point
// DFTs in registers (the
// symmetric cos/sin form for odd p). Prime radices 19..255 run on the matrix
// cores (pass_mfma, in kernel variants of their own), and 257..511 too in
// lean plans (a 512-thread variant); larger primes use an LDS-staged variant
// that computes two outputs per work item.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <algorithm>
#include <cmath>
#include <map>
#include <type_traits>
#include <vector>

#include "zgl_internal.hpp"

// The kernels are compiled in three translation units, one per (arithmetic,
// grid) pair, so that they build in parallel: ZGL_FFT_PART 0 (this file) holds
// the plan functions and the FP64 entry points, zgl_fft_f32.hip (part 1) the
// FP32-arithmetic entry points on FP64 grids, zgl_fft_f32g.hip (part 2) those
// on FP32 grids.
#ifndef ZGL_FFT_PART
#define ZGL_FFT_PART 0
#endif

namespace zgl {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBigItems = 9;
constexpr int kMaxRegRadix = 17;  // larger primes use the LDS-staged pass
constexpr int kBatch = 8;         // global loads in flight per thread
constexpr bool kBwdSpecPairs = false;  // backward: FP32 spectrum read two values per lane

// Threads per workgroup of a kernel variant: 512 for FP32 tiles of 32 column
// pairs in the plain and lean variants (their grid rows are read in 128-byte
// segments, and two workgroups per CU at n = 255 still make 16 waves), 256
// otherwise.
template <int T, class R, int MF>
__host__ __device__ constexpr int fft_threads() {
    return (MF == 3 || (sizeof(R) == 4 && T == 32 && MF != 1)) ? 2 * kThreads : kThreads;
}

template <class R> struct Cplx;
template <> struct Cplx<double> { using type = double2; };
template <> struct Cplx<float> { using type = float2; };

// Vectorised memory phases (VEC kernels, Zernike grids): each lane moves 16
// bytes per global access -- VG = 16 / sizeof(G) consecutive grid columns of a
// latitude row, or VS = 2 consecutive FP32 spectrum values -- so a wave keeps
// 4x (FP32 grids) or 2x (FP64 grids, FP32 spectra) the bytes in flight per
// load instruction of the one-value-per-lane form. The host takes these
// variants when the columns of a row are a multiple of VG and every instance
// base is 16-byte aligned.
template <class G> struct Vec16;
template <> struct Vec16<float> { using type = float4; };
template <> struct Vec16<double> { using type = double2; };
template <class G, bool VEC> constexpr int grid_vec() { return VEC ? 16 / (int)sizeof(G) : 1; }
template <class R, bool VEC> constexpr int spec_vec() { return (VEC && sizeof(R) == 4) ? 2 : 1; }
template <class V> __device__ inline auto vget(const V& v, int k) {
    if constexpr (sizeof(v.x) == 4) return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
    else return k == 0 ? v.x : v.y;
}
template <class V, class S> __device__ inline void vset(V& v, int k, S x) {
    if constexpr (sizeof(v.x) == 4) {
        if (k == 0) v.x = x; else if (k == 1) v.y = x; else if (k == 2) v.z = x; else v.w = x;
    } else {
        if (k == 0) v.x = x; else v.y = x;
    }
}

template <class C> __device__ inline C mk(decltype(C{}.x) x, decltype(C{}.x) y) { C c; c.x = x; c.y = y; return c; }
template <class C> __device__ inline C cadd(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <class C> __device__ inline C csub(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
template <class C> __device__ inline C cmul(C a, C b) {
    return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// In-register p-point DFT, y_s = sum_r x_r e^{DIR 2 pi i r s / p}; `twl` holds
// e^{-2 pi i t / n} and p divides n.
template <int P, int DIR, class C>
__device__ __forceinline__ void dft_small(C (&x)[P], const C* twl, int n) {
    using R = decltype(C{}.x);
    if constexpr (P == 2) {
        const C a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    } else if constexpr (P == 4) {
        const C a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
        const C c = cadd(x[1], x[3]), d = csub(x[1], x[3]);
        const C mid = mk<C>(d.y, -d.x);  // -i d
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
        R cs[H + 1], sn[H + 1];
#pragma unroll
        for (int t = 1; t <= H; ++t) {
            const C w = twl[t * (n / P)];
            cs[t] = w.x;
            sn[t] = -w.y;
        }
        const C x0 = x[0];
        C y0 = x0;
#pragma unroll
        for (int r = 1; r <= H; ++r) {
            const C u = cadd(x[r], x[P - r]);
            const C v = csub(x[r], x[P - r]);
            x[r] = u;
            x[P - r] = v;
            y0 = cadd(y0, u);
        }
        C ys[H + 1], yc[H + 1];
#pragma unroll
        for (int s = 1; s <= H; ++s) {
            C a = x0, b = mk<C>(0, 0);
#pragma unroll
            for (int r = 1; r <= H; ++r) {
                const int t = (r * s) % P;
                const R c = (t <= H) ? cs[t] : cs[P - t];
                const R si = (t <= H) ? sn[t] : -sn[P - t];
                a.x += x[r].x * c;
                a.y += x[r].y * c;
                b.x += x[P - r].x * si;
                b.y += x[P - r].y * si;
            }
            // forward: y_s = a - i b, y_{P-s} = a + i b
            const C mib = mk<C>(b.y, -b.x);
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

// Input twiddle of slot `pos` in a pass; `tw` is that pass's table (or null).
template <int DIR, class C>
__device__ inline C twiddled(C v, const C* tw, int pos) {
    if (!tw) return v;
    C w = tw[pos];
    if (DIR > 0) w.y = -w.y;
    return cmul(v, w);
}

template <int P, int T, int DIR, int NT, class C>
__device__ __forceinline__ void pass_reg(C* data, const C* twl, const C* tw, int n, int L) {
    const int nbf = n / P;
    for (int it = threadIdx.x; it < nbf * T; it += NT) {
        const int t = it % T;
        const int bf = it / T;
        const int bq = bf / L;
        const int q = bf - bq * L;
        const int base = bq * P * L + q;
        C x[P];
#pragma unroll
        for (int r = 0; r < P; ++r) {
            const int pos = base + r * L;
            x[r] = r ? twiddled<DIR>(data[pos * T + t], tw, pos) : data[pos * T + t];
        }
        dft_small<P, DIR>(x, twl, n);
#pragma unroll
        for (int s = 0; s < P; ++s) data[(base + s * L) * T + t] = x[s];
    }
}

// Prime radix P > 17: work item = (butterfly, transform, s in [0, H]) producing
// y_s and y_{P-s}; results are held in registers across a barrier because the
// DFT reads every input slot of its butterfly.
template <int T, int DIR, int NT, class C>
__device__ __forceinline__ void pass_big(C* data, const C* twl, const C* tw, int n, int P, int L) {
    using R = decltype(C{}.x);
    const int H = (P - 1) / 2;
    const int nbf = n / P;
    const int items = nbf * T * (H + 1);
    const int stride = n / P;
    C ya[kMaxBigItems], yb[kMaxBigItems];
#pragma unroll
    for (int c = 0; c < kMaxBigItems; ++c) {
        const int it = threadIdx.x + c * NT;
        if (it >= items) break;
        const int s = it % (H + 1);
        const int rest = it / (H + 1);
        const int t = rest % T;
        const int bf = rest / T;
        const int bq = bf / L;
        const int q = bf - bq * L;
        const int base = bq * P * L + q;
        const C x0 = data[base * T + t];
        if (s == 0) {
            C y = x0;
            for (int r = 1; r < P; ++r) {
                const int pos = base + r * L;
                y = cadd(y, twiddled<DIR>(data[pos * T + t], tw, pos));
            }
            ya[c] = y;
            yb[c] = y;
        } else {
            C a = x0, b = mk<C>(0, 0);
            int idx = 0;
            for (int r = 1; r <= H; ++r) {
                const int pr = base + r * L, pm = base + (P - r) * L;
                const C xr = twiddled<DIR>(data[pr * T + t], tw, pr);
                const C xm = twiddled<DIR>(data[pm * T + t], tw, pm);
                idx += s;
                if (idx >= P) idx -= P;
                const C w = twl[idx * stride];
                const R c0 = w.x, s0 = -w.y;
                a.x += (xr.x + xm.x) * c0;
                a.y += (xr.y + xm.y) * c0;
                b.x += (xr.x - xm.x) * s0;
                b.y += (xr.y - xm.y) * s0;
            }
            const C mib = mk<C>(b.y, -b.x);
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
        const int it = threadIdx.x + c * NT;
        if (it >= items) break;
        const int s = it % (H + 1);
        const int rest = it / (H + 1);
        const int t = rest % T;
        const int bf = rest / T;
        const int bq = bf / L;
        const int q = bf - bq * L;
        const int base = bq * P * L + q;
        data[(base + s * L) * T + t] = ya[c];
        if (s > 0) data[(base + (P - s) * L) * T + t] = yb[c];
    }
}

// Prime radix 17 < P <= kMaxMfmaRadix on the matrix cores. With H = (P-1)/2,
// u_r = x_r + x_{P-r} and v_r = x_r - x_{P-r} (inputs twiddled, u_0 = x_0,
// v_0 = 0), the pass is two real products over r = 0..H,
//   a_s = sum_r cos(2 pi r s / P) u_r,   b_s = sum_r sin(2 pi r s / P) v_r,
// for s = 0..H, and y_s = a_s -/+ i b_s, y_{P-s} = a_s +/- i b_s (the same
// combination as pass_big). Rows s and the contraction index r go to 16x16x4
// MFMA tiles (cos and sin from the LDS twiddle table, index r*s mod P kept
// incrementally); the columns are the (butterfly, sequence) pairs of the
// pass, NC = (n / P) * T. A round computes two tiles per wave (8 or 16) --
// every M block of up to 2 waves / MB column blocks -- holds them in registers
// across a barrier (the pass is in place) and writes them back. Each column's
// arithmetic is the same whatever T, so results do not depend on the tile
// width. The pass runs in its own kernel variants (its accumulators would
// otherwise cost every FFT kernel occupancy).
constexpr int kMfmaTasks = 2;       // tiles per wave and round
constexpr int kMaxMfmaRadix = 255;  // MB = ceil((H + 1) / 16) <= 4 * kMfmaTasks

__host__ __device__ inline bool mfma_radix(int P) { return P > kMaxRegRadix && P <= kMaxMfmaRadix; }
// Primes up to 511 (MB <= 16) fit the pass on eight waves: the 512-thread
// lean variant (MF = 3) of the widths wide_mf gives it.
constexpr int kMaxWideRadix = 511;
__host__ __device__ inline bool wide_radix(int P) { return P > kMaxMfmaRadix && P <= kMaxWideRadix; }
template <int T>
constexpr int wide_mf() { return (T == 4 || T == 8 || T == 16) ? 3 : 1; }

template <class R> struct PassMfma;
template <> struct PassMfma<double> {
    typedef double acc __attribute__((ext_vector_type(4)));
    __device__ static acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // lane holds D[lane/16 + 4r][lane%16]
    __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct PassMfma<float> {
    typedef float acc __attribute__((ext_vector_type(4)));
    __device__ static acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // lane holds D[4(lane/16) + r][lane%16]
    __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <int T, int DIR, int NT, class C>
__device__ __forceinline__ void pass_mfma(C* data, const C* twl, const C* tw, int n, int P, int L) {
    constexpr int NW = NT / 64;  // waves
    using R = decltype(C{}.x);
    using M = PassMfma<R>;
    using A = typename M::acc;
    const int H = (P - 1) / 2;
    const int NC = (n / P) * T;
    const int MB = (H + 16) >> 4, NB = (NC + 15) >> 4;
    const int NBR = max(1, NW * kMfmaTasks / MB);  // column blocks per round
    const int KS = (H + 4) >> 2;       // k-steps over r = 0..H
    const int stride = n / P;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, g = lane >> 4;

    for (int nb0 = 0; nb0 < NB; nb0 += NBR) {
        const int ntask = MB * min(NBR, NB - nb0);
        A ar[kMfmaTasks], ai[kMfmaTasks], br[kMfmaTasks], bi[kMfmaTasks];
        int sA[kMfmaTasks], idx[kMfmaTasks], step[kMfmaTasks], base[kMfmaTasks], tcol[kMfmaTasks];
        bool cok[kMfmaTasks];
#pragma unroll
        for (int q = 0; q < kMfmaTasks; ++q) {
            ar[q] = ai[q] = br[q] = bi[q] = A{0, 0, 0, 0};
            const int task = wave + NW * q;
            const int mb = task % MB, nb = nb0 + task / MB;
            // A operand of this lane: row s, contraction index r = 4 kk + g
            sA[q] = 16 * mb + i;
            step[q] = (4 * sA[q]) % P;
            idx[q] = (g * sA[q]) % P;
            // B operand of this lane: row r = 4 kk + g, column c
            const int c = 16 * nb + i;
            cok[q] = task < ntask && c < NC;
            const int bf = c / T;
            tcol[q] = c - bf * T;
            const int bq = bf / L;
            base[q] = bq * P * L + (bf - bq * L);
        }
        for (int kk = 0; kk < KS; ++kk) {
            const int r = 4 * kk + g;
#pragma unroll
            for (int q = 0; q < kMfmaTasks; ++q) {
                if (wave + NW * q >= ntask) break;  // uniform in the wave
                R ca = 0, sa = 0;
                if (sA[q] <= H && r <= H) {
                    const C w = twl[idx[q] * stride];
                    ca = w.x;
                    sa = -w.y;
                }
                idx[q] += step[q];
                if (idx[q] >= P) idx[q] -= P;
                C u = mk<C>(0, 0), v = mk<C>(0, 0);
                if (cok[q] && r <= H) {
                    const int t = tcol[q];
                    if (r == 0) {
                        u = data[base[q] * T + t];
                    } else {
                        const int pr = base[q] + r * L, pm = base[q] + (P - r) * L;
                        const C xr = twiddled<DIR>(data[pr * T + t], tw, pr);
                        const C xm = twiddled<DIR>(data[pm * T + t], tw, pm);
                        u = cadd(xr, xm);
                        v = csub(xr, xm);
                    }
                }
                ar[q] = M::mma(ca, u.x, ar[q]);
                ai[q] = M::mma(ca, u.y, ai[q]);
                br[q] = M::mma(sa, v.x, br[q]);
                bi[q] = M::mma(sa, v.y, bi[q]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kMfmaTasks; ++q) {
            const int task = wave + NW * q;
            if (task >= ntask) break;
            const int mb = task % MB, nb = nb0 + task / MB;
            const int c = 16 * nb + i;  // output column of this lane
            if (c >= NC) continue;
            const int bf = c / T, t = c - bf * T;
            const int bq = bf / L;
            const int b0 = bq * P * L + (bf - bq * L);
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int s = 16 * mb + M::row(lane, rr);
                if (s > H) continue;
                const C a = mk<C>(ar[q][rr], ai[q][rr]);
                const C mib = mk<C>(bi[q][rr], -br[q][rr]);  // -i b
                C ya, yb;
                if (DIR < 0) {
                    ya = cadd(a, mib);
                    yb = csub(a, mib);
                } else {
                    ya = csub(a, mib);
                    yb = cadd(a, mib);
                }
                data[(b0 + s * L) * T + t] = ya;
                if (s > 0) data[(b0 + (P - s) * L) * T + t] = yb;
            }
        }
        __syncthreads();
    }
}

// MF: 0, the kernel variant without the matrix-core pass (prime radices
// above kMaxRegRadix take pass_big); 1, with it; 2, the lean variant for
// plans whose other radices are at most kLeanRadix (no 11-, 13- or 17-point
// register DFT compiled in, so fewer registers and more waves per SIMD).
constexpr int kLeanRadix = 7;

template <int T, int DIR, int MF, int NT, class C>
__device__ __forceinline__ void run_passes(C* data, const C* twl, const C* twp, const FftDesc& d) {
    const int n = d.n;
    for (int ps = 0; ps < d.npass; ++ps) {
        __syncthreads();
        const int P = d.radix[ps], L = d.L[ps];
        const C* tw = ps ? twp + (ps - 1) * n : nullptr;
        if constexpr (MF >= 2) {
            switch (P) {
                case 2: pass_reg<2, T, DIR, NT>(data, twl, tw, n, L); break;
                case 3: pass_reg<3, T, DIR, NT>(data, twl, tw, n, L); break;
                case 4: pass_reg<4, T, DIR, NT>(data, twl, tw, n, L); break;
                case 5: pass_reg<5, T, DIR, NT>(data, twl, tw, n, L); break;
                case 7: pass_reg<7, T, DIR, NT>(data, twl, tw, n, L); break;
                default: pass_mfma<T, DIR, NT>(data, twl, tw, n, P, L); break;
            }
            continue;
        }
        switch (P) {
            case 2: pass_reg<2, T, DIR, NT>(data, twl, tw, n, L); break;
            case 3: pass_reg<3, T, DIR, NT>(data, twl, tw, n, L); break;
            case 4: pass_reg<4, T, DIR, NT>(data, twl, tw, n, L); break;
            case 5: pass_reg<5, T, DIR, NT>(data, twl, tw, n, L); break;
            case 7: pass_reg<7, T, DIR, NT>(data, twl, tw, n, L); break;
            case 11: pass_reg<11, T, DIR, NT>(data, twl, tw, n, L); break;
            case 13: pass_reg<13, T, DIR, NT>(data, twl, tw, n, L); break;
            case 17: pass_reg<17, T, DIR, NT>(data, twl, tw, n, L); break;
            default:
                if constexpr (MF) {
                    if (mfma_radix(P)) {
                        pass_mfma<T, DIR, NT>(data, twl, tw, n, P, L);
                        break;
                    }
                }
                pass_big<T, DIR, NT>(data, twl, tw, n, P, L);
                break;
        }
    }
    __syncthreads();
}

struct PairMap {
    int ncol, nth, jh, lat_stride, lon_stride;
    int ntiles;  // tiles per instance
};

// Workgroups are dealt round-robin to the 8 XCDs (workgroup b runs on XCD
// b % 8), each with its own L2. The grid has 8 * ceil(ntiles / 8) workgroups
// per instance and workgroup b takes tile (b % 8) * q + b / 8, so every XCD
// walks a contiguous range of tiles: neighbouring column tiles, which share
// cache lines at their edges, meet in the same L2.
__device__ inline int xcd_tile(const PairMap& pm) {
    if (pm.ntiles < 8) return blockIdx.x;  // one workgroup per tile (see grid_tiles)
    const int q = (pm.ntiles + 7) >> 3;
    const int b = blockIdx.x;
    return (b & 7) * q + (b >> 3);
}

// A tile holds T consecutive columns f = h * ncol + c of the flattened
// (latitude pair h, radial node c) index, so a tile may end one pair and
// start the next and every tile but the last is full (SH grids: ncol = 1,
// f = h). Every column a thread touches is threadIdx.x % T (the thread count
// NT is a multiple of T), so each thread resolves its (h, c) once; h >= jh
// marks a column past the end.
template <int T, int NT>
__device__ inline void thread_column(const PairMap& pm, int tile, int& h, int& c) {
    static_assert(NT % T == 0, "a thread keeps one column");
    const int f = tile * T + (int)threadIdx.x % T;
    h = f / pm.ncol;
    c = f - h * pm.ncol;
}

// Narrow tiles keep the twiddle tables in global
// memory (read through L1/L2, where every workgroup finds the same few KB);
// wider tiles stage them in LDS. Only long sequences take such narrow tiles,
// and beside them the tables would take a third of the LDS or more and cut
// the workgroups per CU (n = 1023, FP64, T = 4: 50 KB of tables beside 65 KB
// of data). FP64 keeps them global up to T = 8 (n = 511: 66 KB of data
// alone, two workgroups per CU only without the tables); FP32 only up to
// T = 4 (n = 1023, T = 8: the forward pass measured 15% slower with global
// tables than with them in LDS). Fixed per width and element size, so the compiler
// knows each table's address space (LDS or global loads, never flat ones).
__host__ __device__ constexpr bool twiddles_in_lds(int T, size_t csize) {
    return T > (csize == 16 ? 8 : 4);
}

// copy = false: the LDS already holds these tables (a persistent workgroup's
// previous item was of the same group); only the pointers are set.
template <int T, int NT, class C>
__device__ inline void stage_tables(C*& data, const C*& twl, const C*& twp, unsigned short*& perm,
                                    const FftDesc& d, const C* twg, const C* twpg,
                                    const unsigned short* permg, bool copy = true) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n = d.n;
    data = reinterpret_cast<C*>(smem_raw);
    if constexpr (twiddles_in_lds(T, sizeof(C))) {
        C* tl = data + n * T;
        C* tp = tl + n;
        perm = reinterpret_cast<unsigned short*>(tp + (d.npass > 0 ? d.npass - 1 : 0) * n);
        if (copy) {
            for (int i = threadIdx.x; i < n; i += NT) tl[i] = twg[i];
            for (int i = threadIdx.x; i < (d.npass - 1) * n; i += NT) tp[i] = twpg[i];
        }
        twl = tl;
        twp = tp;
    } else {
        perm = reinterpret_cast<unsigned short*>(data + n * T);
        twl = twg;
        twp = twpg;
    }
    if (copy)
        for (int i = threadIdx.x; i < n; i += NT) perm[i] = permg[i];
    __syncthreads();
}

// Waves per SIMD asked of the lean variant (3 in FP64, 4 in FP32), of the
// 512-thread one (4: two workgroups per CU) and of the narrow FP64 tiles with
// the matrix-core pass (T <= 8, long sequences such as SH L = 512: 2); none
// of the others. Unasked, the compiler gives those narrow kernels one
// register more than two waves allow (a wave per SIMD less: SH L = 512 FFTs
// 4.3 -> 7.5 ms per 1024 instances). Asking three of the narrow FP64 kernels
// without the pass makes them spill.
template <int T, class R, int MF>
constexpr int lean_waves() {
    if constexpr (sizeof(R) == 8 && T <= 8 && MF == 1) return 2;
    return MF == 2 ? (sizeof(R) == 8 ? 3 : 4) : (fft_threads<T, R, MF>() > kThreads ? (sizeof(R) == 8 ? 2 : 4) : 1);
}

// The arguments of one FFT launch over the instances of one order (also one
// group of a persistent launch).
template <class R, class G>
struct FftFwdArgs {
    FftDesc d;
    const typename Cplx<R>::type* twg;
    const typename Cplx<R>::type* twpg;
    const unsigned short* permg;
    Inst<const G> grid;
    typename Cplx<R>::type* spec;
    long spec_stride;
    PairMap pm;
    int nm;
    double scale;
    const double* wlat;
    const double* wrad;
    double center_factor;
};

template <class R, class G>
struct FftBwdArgs {
    FftDesc d;
    const typename Cplx<R>::type* twg;
    const typename Cplx<R>::type* twpg;
    const unsigned short* permg;
    const typename Cplx<R>::type* spec;
    long spec_stride;
    Inst<G> grid;
    PairMap pm;
    int nm;
    double scale;
    const double* wlat;
    const double* wrad;
    double center_factor;
};

template <int T, class R, class G, int NT, bool VEC>
__device__ __forceinline__ void fft_forward_output(const FftFwdArgs<R, G>& a, int tile, int inst,
                                                   const typename Cplx<R>::type* data);

// One tile of one instance; copy_tables = false when the LDS already holds
// this order's tables.
template <int T, int LK, class R, class G, int MF, bool VEC>
__device__ __forceinline__ void fft_forward_body(const FftFwdArgs<R, G>& a, const int tile, const int inst,
                                                 const bool copy_tables) {
    using C = typename Cplx<R>::type;
    constexpr int NT = fft_threads<T, R, MF>();
    const FftDesc& d = a.d;
    const PairMap& pm = a.pm;
    C* data;
    const C *twl, *twp;
    unsigned short* perm;
    // the tables and operands through restrict-qualified locals: no store of
    // the kernel aliases them (without this the FP64 kernels hold ~30 more
    // VGPRs and lose a wave per SIMD)
    const C* __restrict__ twg = a.twg;
    const C* __restrict__ twpg = a.twpg;
    const unsigned short* __restrict__ permg = a.permg;
    stage_tables<T, NT>(data, twl, twp, perm, d, twg, twpg, permg, copy_tables);
    const int n = d.n;
    int th, tc;
    thread_column<T, NT>(pm, tile, th, tc);
    const int h0 = tile * T;  // first latitude pair of a row-layout (SH) tile
    const G* __restrict__ g = a.grid.at(inst);

    // Loads are issued in batches of kBatch per thread before any LDS store,
    // so each thread keeps that many independent global reads in flight.
    if constexpr (VEC) {
        // VG columns per lane; the lane's column group is threadIdx.x % (T / VG).
        // KB loads of each latitude in flight per thread: on FP64 grids with
        // FP64 arithmetic, or with 256 threads, a whole tile at n = 255, T = 16
        // (8 per thread) goes out in one batch, one memory round trip instead
        // of two
        constexpr int VG = grid_vec<G, VEC>(), TQ = T / VG;
        constexpr int KB = (VG == 2 && (sizeof(R) == 8 || NT == kThreads)) ? kBatch : kBatch / 2;
        using V = typename Vec16<G>::type;
        static_assert(LK == GRID_COLS && T % VG == 0 && NT % TQ == 0, "vector tile");
        const int f = tile * T + ((int)threadIdx.x % TQ) * VG;
        const int h = f / pm.ncol, c = f - h * pm.ncol;
        const int jn = pm.nth / 2 + h, js = pm.jh - 1 - h;
        const G* col = g + c;
        for (int e0 = threadIdx.x; e0 < n * TQ; e0 += NT * KB) {
            V av[KB], bv[KB];
#pragma unroll
            for (int u = 0; u < KB; ++u) {
                const int e = e0 + u * NT;
                const int i = e / TQ;
                av[u] = V{};
                bv[u] = V{};
                if (e < n * TQ && h < pm.jh) {
                    const G* row = col + (long)i * pm.lon_stride;
                    av[u] = *reinterpret_cast<const V*>(row + (long)jn * pm.lat_stride);
                    if (js != jn) bv[u] = *reinterpret_cast<const V*>(row + (long)js * pm.lat_stride);
                }
            }
#pragma unroll
            for (int u = 0; u < KB; ++u) {
                const int e = e0 + u * NT;
                if (e < n * TQ) {
                    C* dst = data + perm[e / TQ] * T + (e % TQ) * VG;
#pragma unroll
                    for (int k = 0; k < VG; ++k) dst[k] = mk<C>((R)vget(av[u], k), (R)vget(bv[u], k));
                }
            }
        }
    } else if (LK == GRID_COLS) {
        for (int e0 = threadIdx.x; e0 < n * T; e0 += NT * kBatch) {
            R av[kBatch], bv[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int e = e0 + u * NT;
                const int i = e / T, h = th, c = tc;
                av[u] = 0;
                bv[u] = 0;
                if (e < n * T && h < pm.jh) {
                    const int jn = pm.nth / 2 + h, js = pm.jh - 1 - h;
                    const G* row = g + (long)i * pm.lon_stride + c;
                    av[u] = (R)row[(long)jn * pm.lat_stride];
                    if (js != jn) bv[u] = (R)row[(long)js * pm.lat_stride];
                }
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int e = e0 + u * NT;
                if (e < n * T) data[perm[e / T] * T + e % T] = mk<C>(av[u], bv[u]);
            }
        }
    } else {
        for (int e0 = threadIdx.x; e0 < n * T; e0 += NT * kBatch) {
            R av[kBatch], bv[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int e = e0 + u * NT;
                const int t = e / n, i = e - t * n;
                const int h = h0 + t;
                av[u] = 0;
                bv[u] = 0;
                if (e < n * T && h < pm.jh) {
                    const int jn = pm.nth / 2 + h, js = pm.jh - 1 - h;
                    av[u] = (R)g[(long)jn * pm.lat_stride + i];
                    if (js != jn) bv[u] = (R)g[(long)js * pm.lat_stride + i];
                }
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int e = e0 + u * NT;
                if (e < n * T) {
                    const int t = e / n, i = e - t * n;
                    data[perm[i] * T + t] = mk<C>(av[u], bv[u]);
                }
            }
        }
    }

    run_passes<T, -1, MF, NT>(data, twl, twp, d);
    fft_forward_output<T, R, G, NT, VEC>(a, tile, inst, data);
}

// The output phase of the forward FFT: the parity split, weights and the
// spectrum stores of one tile, from the transformed sequences in LDS.
template <int T, class R, class G, int NT, bool VEC>
__device__ __forceinline__ void fft_forward_output(const FftFwdArgs<R, G>& a, const int tile, const int inst,
                                                   const typename Cplx<R>::type* data) {
    using C = typename Cplx<R>::type;
    const PairMap& pm = a.pm;
    const int n = a.d.n;
    const int nm = a.nm;
    const double scale = a.scale, center_factor = a.center_factor;
    const double* __restrict__ wlat = a.wlat;
    const double* __restrict__ wrad = a.wrad;
    C* __restrict__ sp = a.spec + (long)inst * a.spec_stride;
    int th, tc;
    thread_column<T, NT>(pm, tile, th, tc);
    if constexpr (spec_vec<R, VEC>() == 2) {
        // two consecutive columns per lane: 16-byte spectrum stores
        constexpr int TP = T / 2;
        static_assert(T % 2 == 0 && NT % TP == 0, "column pairs");
        const int f = tile * T + ((int)threadIdx.x % TP) * 2;
        const int h = f / pm.ncol, c = f - h * pm.ncol;
        if (h < pm.jh) {
            const bool mid = pm.nth / 2 + h == pm.jh - 1 - h;
            // the same expression as the one-column form, for identical roundings
            const R W0 = (R)((wlat ? wlat[h] * (wrad ? wrad[c] : 1.0) : 1.0) * scale);
            const R W1 = (R)((wlat ? wlat[h] * (wrad ? wrad[c + 1] : 1.0) : 1.0) * scale);
            const R cf = (R)center_factor;
            float4* sp4 = reinterpret_cast<float4*>(sp);
            for (int e = threadIdx.x; e < nm * TP; e += NT) {
                const int m = e / TP, t = (e % TP) * 2;
                C E[2], O[2];
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const C zm = data[m * T + t + k];
                    const C zc = data[(m ? n - m : 0) * T + t + k];
                    const C A = mk<C>(R(0.5) * (zm.x + zc.x), R(0.5) * (zm.y - zc.y));
                    const C B = mk<C>(R(0.5) * (zm.y + zc.y), R(-0.5) * (zm.x - zc.x));
                    const R W = k ? W1 : W0;
                    const C Ha = mk<C>(W * A.x, -W * A.y);
                    const C Hb = mk<C>(W * B.x, -W * B.y);
                    if (mid) {
                        E[k] = mk<C>(cf * Ha.x, cf * Ha.y);
                        O[k] = mk<C>(0, 0);
                    } else {
                        E[k] = cadd(Ha, Hb);
                        O[k] = csub(Ha, Hb);
                    }
                }
                const long fe = ((long)(m * 2 + 0) * pm.jh + h) * pm.ncol + c;
                const long fo = ((long)(m * 2 + 1) * pm.jh + h) * pm.ncol + c;
                sp4[fe / 2] = make_float4(E[0].x, E[0].y, E[1].x, E[1].y);
                sp4[fo / 2] = make_float4(O[0].x, O[0].y, O[1].x, O[1].y);
            }
        }
        return;
    }
    // a thread keeps one column (h, c) for every m: its quadrature weight
    // (none for the transpose of the backward transform) is read once, not
    // per m (a global load the compiler cannot hoist out of the guarded loop)
    const int h = th, c = tc;
    if (h >= pm.jh) return;
    const R W = (R)((wlat ? wlat[h] * (wrad ? wrad[c] : 1.0) : 1.0) * scale);
    for (int e = threadIdx.x; e < nm * T; e += NT) {
        const int m = e / T, t = e % T;
        const C zm = data[m * T + t];
        const C zc = data[(m ? n - m : 0) * T + t];
        // two-for-one split: A = (Z_m + conj Z_{n-m}) / 2, B = (Z_m - conj Z_{n-m}) / (2i)
        const C A = mk<C>(R(0.5) * (zm.x + zc.x), R(0.5) * (zm.y - zc.y));
        const C B = mk<C>(R(0.5) * (zm.y + zc.y), R(-0.5) * (zm.x - zc.x));
        // H = weight * conj(scale * X)
        const C Ha = mk<C>(W * A.x, -W * A.y);
        const C Hb = mk<C>(W * B.x, -W * B.y);
        C E, O;
        if (pm.nth / 2 + h == pm.jh - 1 - h) {
            const R cf = (R)center_factor;
            E = mk<C>(cf * Ha.x, cf * Ha.y);
            O = mk<C>(0, 0);
        } else {
            E = cadd(Ha, Hb);
            O = csub(Ha, Hb);
        }
        sp[((long)(m * 2 + 0) * pm.jh + h) * pm.ncol + c] = E;
        sp[((long)(m * 2 + 1) * pm.jh + h) * pm.ncol + c] = O;
    }
}

template <int T, int LK, class R, class G, int MF, bool VEC>
__device__ __forceinline__ void fft_backward_body(const FftBwdArgs<R, G>& a, const int tile, const int inst,
                                                  const bool copy_tables) {
    using C = typename Cplx<R>::type;
    constexpr int NT = fft_threads<T, R, MF>();
    const FftDesc& d = a.d;
    const PairMap& pm = a.pm;
    const int nm = a.nm;
    const double scale = a.scale, center_factor = a.center_factor;
    const double* __restrict__ wlat = a.wlat;
    const double* __restrict__ wrad = a.wrad;
    C* data;
    const C *twl, *twp;
    unsigned short* perm;
    // the tables and operands through restrict-qualified locals: no store of
    // the kernel aliases them (without this the FP64 kernels hold ~30 more
    // VGPRs and lose a wave per SIMD)
    const C* __restrict__ twg = a.twg;
    const C* __restrict__ twpg = a.twpg;
    const unsigned short* __restrict__ permg = a.permg;
    stage_tables<T, NT>(data, twl, twp, perm, d, twg, twpg, permg, copy_tables);
    const int n = d.n;
    int th, tc;
    thread_column<T, NT>(pm, tile, th, tc);
    const int h0 = tile * T;  // first latitude pair of a row-layout (SH) tile
    G* __restrict__ g = a.grid.at(inst);
    const C* __restrict__ sp = a.spec + (long)inst * a.spec_stride;
    const int F = n / 2 + 1;
    if constexpr (spec_vec<R, VEC>() == 2 && kBwdSpecPairs) {
        // two consecutive columns per lane: 16-byte spectrum loads
        constexpr int TP = T / 2;
        static_assert(T % 2 == 0 && NT % TP == 0, "column pairs");
        const int f = tile * T + ((int)threadIdx.x % TP) * 2;
        const int h = f / pm.ncol, c = f - h * pm.ncol;
        const bool mid = pm.nth / 2 + h == pm.jh - 1 - h;
        const float4* sp4 = reinterpret_cast<const float4*>(sp);
        for (int e0 = threadIdx.x; e0 < F * TP; e0 += NT * kBatch) {
            float4 ae[kBatch], ao[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int e = e0 + u * NT;
                const int m = e / TP;
                ae[u] = make_float4(0, 0, 0, 0);
                ao[u] = make_float4(0, 0, 0, 0);
                if (e < F * TP && h < pm.jh && m < nm) {
                    ae[u] = sp4[(((long)(m * 2 + 0) * pm.jh + h) * pm.ncol + c) / 2];
                    ao[u] = sp4[(((long)(m * 2 + 1) * pm.jh + h) * pm.ncol + c) / 2];
                }
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int e = e0 + u * NT;
                if (e >= F * TP) continue;
                const int m = e / TP, t0 = (e % TP) * 2;
                const R mu = m > 0 ? R(0.5) : R(1);
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const C a = k ? mk<C>(ae[u].z, ae[u].w) : mk<C>(ae[u].x, ae[u].y);
                    const C o = k ? mk<C>(ao[u].z, ao[u].w) : mk<C>(ao[u].x, ao[u].y);
                    C Sa, Sb;
                    if (mid) {
                        Sa = mk<C>(mu * a.x, -mu * a.y);
                        Sb = mk<C>(0, 0);
                    } else {
                        const C Ev = cadd(a, o), Dv = csub(a, o);
                        Sa = mk<C>(mu * Ev.x, -mu * Ev.y);
                        Sb = mk<C>(mu * Dv.x, -mu * Dv.y);
                    }
                    C zm, zp = mk<C>(0, 0);
                    if (m == 0 || 2 * m == n) {
                        zm = mk<C>(Sa.x, Sb.x);
                    } else {
                        zm = mk<C>(Sa.x - Sb.y, Sa.y + Sb.x);
                        zp = mk<C>(Sa.x + Sb.y, -Sa.y + Sb.x);
                    }
                    data[perm[m] * T + t0 + k] = zm;
                    if (m > 0 && 2 * m != n) data[perm[n - m] * T + t0 + k] = zp;
                }
            }
        }
    } else {
    for (int e0 = threadIdx.x; e0 < F * T; e0 += NT * kBatch) {
        C ae[kBatch], ao[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int e = e0 + u * NT;
            const int m = e / T, h = th, c = tc;
            ae[u] = mk<C>(0, 0);
            ao[u] = mk<C>(0, 0);
            if (e < F * T && h < pm.jh && m < nm) {
                ae[u] = sp[((long)(m * 2 + 0) * pm.jh + h) * pm.ncol + c];
                ao[u] = sp[((long)(m * 2 + 1) * pm.jh + h) * pm.ncol + c];
            }
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int e = e0 + u * NT;
            if (e >= F * T) continue;
            const int m = e / T, t = e % T, h = th;
            const R mu = m > 0 ? R(0.5) : R(1);
            C Sa, Sb;
            if (pm.nth / 2 + h == pm.jh - 1 - h) {
                Sa = mk<C>(mu * ae[u].x, -mu * ae[u].y);
                Sb = mk<C>(0, 0);
            } else {
                const C Ev = cadd(ae[u], ao[u]), Dv = csub(ae[u], ao[u]);
                Sa = mk<C>(mu * Ev.x, -mu * Ev.y);
                Sb = mk<C>(mu * Dv.x, -mu * Dv.y);
            }
            C zm, zp = mk<C>(0, 0);
            if (m == 0 || 2 * m == n) {
                zm = mk<C>(Sa.x, Sb.x);  // imaginary parts of X_0 (and Nyquist) ignored
            } else {
                zm = mk<C>(Sa.x - Sb.y, Sa.y + Sb.x);   // Sa + i Sb
                zp = mk<C>(Sa.x + Sb.y, -Sa.y + Sb.x);  // conj(Sa) + i conj(Sb)
            }
            data[perm[m] * T + t] = zm;
            if (m > 0 && 2 * m != n) data[perm[n - m] * T + t] = zp;
        }
    }
    }

    run_passes<T, +1, MF, NT>(data, twl, twp, d);

    if constexpr (VEC) {
        // VG columns per lane: 16-byte grid stores
        constexpr int VG = grid_vec<G, VEC>(), TQ = T / VG;
        using V = typename Vec16<G>::type;
        static_assert(LK == GRID_COLS && T % VG == 0 && NT % TQ == 0, "vector tile");
        const int f = tile * T + ((int)threadIdx.x % TQ) * VG;
        const int h = f / pm.ncol, c = f - h * pm.ncol;
        if (h < pm.jh) {
            const int jn = pm.nth / 2 + h, js = pm.jh - 1 - h;
            R W[VG];
#pragma unroll
            for (int k = 0; k < VG; ++k) {
                W[k] = 1;
                if (wlat) {
                    W[k] = (R)(wlat[h] * (wrad ? wrad[c + k] : 1.0) * scale);
                    if (jn == js) W[k] *= (R)center_factor;
                }
            }
            G* col = g + c;
            for (int e = threadIdx.x; e < n * TQ; e += NT) {
                const int i = e / TQ, t0 = (e % TQ) * VG;
                V a, b;
#pragma unroll
                for (int k = 0; k < VG; ++k) {
                    C z = data[i * T + t0 + k];
                    if (wlat) {
                        z.x *= W[k];
                        z.y *= W[k];
                    }
                    vset(a, k, (G)z.x);
                    vset(b, k, (G)z.y);
                }
                G* row = col + (long)i * pm.lon_stride;
                *reinterpret_cast<V*>(row + (long)jn * pm.lat_stride) = a;
                if (js != jn) *reinterpret_cast<V*>(row + (long)js * pm.lat_stride) = b;
            }
        }
    } else if (LK == GRID_COLS) {
        // one column (h, c) per thread: its output weight (the transpose of
        // the forward transform) is read once
        const int h = th, c = tc;
        if (h < pm.jh) {
            const int jn = pm.nth / 2 + h, js = pm.jh - 1 - h;
            R W = 1;
            if (wlat) {
                W = (R)(wlat[h] * (wrad ? wrad[c] : 1.0) * scale);
                if (jn == js) W *= (R)center_factor;
            }
            for (int e = threadIdx.x; e < n * T; e += NT) {
                const int i = e / T, t = e % T;
                C z = data[i * T + t];
                if (wlat) {
                    z.x *= W;
                    z.y *= W;
                }
                G* row = g + (long)i * pm.lon_stride + c;
                row[(long)jn * pm.lat_stride] = (G)z.x;
                if (js != jn) row[(long)js * pm.lat_stride] = (G)z.y;
            }
        }
    } else {
        for (int e = threadIdx.x; e < n * T; e += NT) {
            const int t = e / n, i = e - t * n;
            const int h = h0 + t;
            if (h >= pm.jh) continue;
            C z = data[i * T + t];
            const int jn = pm.nth / 2 + h, js = pm.jh - 1 - h;
            if (wlat) {
                R W = (R)(wlat[h] * scale);
                if (jn == js) W *= (R)center_factor;
                z.x *= W;
                z.y *= W;
            }
            g[(long)jn * pm.lat_stride + i] = (G)z.x;
            if (js != jn) g[(long)js * pm.lat_stride + i] = (G)z.y;
        }
    }
}

template <int T, int LK, class R, class G, int MF, bool VEC>
__global__ __launch_bounds__((fft_threads<T, R, MF>())) __attribute__((amdgpu_waves_per_eu(lean_waves<T, R, MF>())))
void fft_forward_kernel(FftFwdArgs<R, G> a) {
    const int tile = xcd_tile(a.pm);
    if (tile >= a.pm.ntiles) return;
    fft_forward_body<T, LK, R, G, MF, VEC>(a, tile, blockIdx.y, true);
}

template <int T, int LK, class R, class G, int MF, bool VEC>
__global__ __launch_bounds__((fft_threads<T, R, MF>())) __attribute__((amdgpu_waves_per_eu(lean_waves<T, R, MF>())))
void fft_backward_kernel(FftBwdArgs<R, G> a) {
    const int tile = xcd_tile(a.pm);
    if (tile >= a.pm.ntiles) return;
    fft_backward_body<T, LK, R, G, MF, VEC>(a, tile, blockIdx.y, true);
}

// Persistent forms (ragged batches of many orders): workgroup b walks the
// items [bounds[b], bounds[b + 1]), item = (group, tile, instance), with the
// arguments of each group in groups[]. Consecutive items of one group reuse
// the tables already in LDS.
template <int T, int LK, class R, class G, int MF, bool VEC>
__global__ __launch_bounds__((fft_threads<T, R, MF>())) __attribute__((amdgpu_waves_per_eu(lean_waves<T, R, MF>())))
void fft_forward_persistent(const FftFwdArgs<R, G>* __restrict__ groups, const int4* __restrict__ items,
                            const int* __restrict__ bounds) {
    const int w1 = bounds[blockIdx.x + 1];
    int prev = -1;
    for (int w = bounds[blockIdx.x]; w < w1; ++w) {
        const int4 it = items[w];
        __syncthreads();  // the previous item is done with the LDS
        fft_forward_body<T, LK, R, G, MF, VEC>(groups[it.x], it.y, it.z, it.x != prev);
        prev = it.x;
    }
}

template <int T, int LK, class R, class G, int MF, bool VEC>
__global__ __launch_bounds__((fft_threads<T, R, MF>())) __attribute__((amdgpu_waves_per_eu(lean_waves<T, R, MF>())))
void fft_backward_persistent(const FftBwdArgs<R, G>* __restrict__ groups, const int4* __restrict__ items,
                             const int* __restrict__ bounds) {
    const int w1 = bounds[blockIdx.x + 1];
    int prev = -1;
    for (int w = bounds[blockIdx.x]; w < w1; ++w) {
        const int4 it = items[w];
        __syncthreads();
        fft_backward_body<T, LK, R, G, MF, VEC>(groups[it.x], it.y, it.z, it.x != prev);
        prev = it.x;
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

size_t lds_bytes(int n, int npass, int T, size_t csize) {
    const size_t tables = twiddles_in_lds(T, csize) ? (size_t)(n + (npass > 0 ? npass - 1 : 0) * n) * csize : 0;
    return (size_t)n * T * csize + tables + 2 * (size_t)n;
}

PairMap pair_map(const Geom& g, int T) {
    PairMap pm;
    pm.ncol = g.ncol;
    pm.nth = g.nth;
    pm.jh = g.jh;
    pm.lat_stride = g.lat_stride;
    pm.lon_stride = g.lon_stride;
    pm.ntiles = (int)(((long)pm.jh * pm.ncol + T - 1) / T);
    return pm;
}

// Workgroups per instance: a multiple of 8 (see xcd_tile); instances of fewer
// than 8 tiles (short SH grids) get one workgroup per tile, not 8 of which
// most would exit at once.
int grid_tiles(const PairMap& pm) { return pm.ntiles < 8 ? pm.ntiles : (pm.ntiles + 7) / 8 * 8; }

template <class R>
const typename Cplx<R>::type* tw_of(const FftPlan& f) {
    if constexpr (sizeof(R) == 8) return f.tw;
    else return f.tw32;
}
template <class R>
const typename Cplx<R>::type* twp_of(const FftPlan& f) {
    if constexpr (sizeof(R) == 8) return f.twp;
    else return f.twp32;
}

// The lean variant is instantiated for the widths prime lengths use (others
// fall back to the general one, which computes the same values).
template <int T>
constexpr int lean_mf() { return (T == 8 || T == 16 || T == 32) ? 2 : 1; }

template <class K>
K pick_variant(const FftPlan& f, K none, K general, K lean, K wide) {
    if (!f.mfma_pass) return none;
    if (f.wide) return wide;
    return f.lean ? lean : general;
}

// Whether the 16-byte-per-lane variant applies (see Vec16): Zernike-style
// column grids, tiles of 8 or more columns, rows a whole number of vectors,
// aligned instance bases, and an even FP32 spectrum stride.
template <int T, class R, class G>
bool use_vec(const FftLaunch& L, const PairMap& pm, const void* spec, long ss) {
    if constexpr (T < 8) {
        return false;
    } else {
        if (L.layout_kind != GRID_COLS || !L.grid_aligned16 || std::getenv("ZGL_FFT_SCALAR")) return false;
        if (pm.ncol % grid_vec<G, true>() != 0) return false;
        if (spec_vec<R, true>() == 2 && (reinterpret_cast<uintptr_t>(spec) % 16 != 0 || ss % 2 != 0)) return false;
        return true;
    }
}

template <int T, class R, class G>
void launch_forward(const FftLaunch& L, Inst<const G> grid, void* spec, long ss, int nb,
                    hipStream_t s) {
    using C = typename Cplx<R>::type;
    const FftPlan& f = L.t->fft;
    const PairMap pm = pair_map(*L.g, T);
    const size_t lds = lds_bytes(f.desc.n, f.desc.npass, T, sizeof(C));
    dim3 grd(grid_tiles(pm), nb);
    auto k = L.layout_kind == GRID_COLS ? pick_variant(f, fft_forward_kernel<T, GRID_COLS, R, G, 0, false>,
                                                      fft_forward_kernel<T, GRID_COLS, R, G, 1, false>,
                                                      fft_forward_kernel<T, GRID_COLS, R, G, lean_mf<T>(), false>,
                                                      fft_forward_kernel<T, GRID_COLS, R, G, wide_mf<T>(), false>)
                                        : pick_variant(f, fft_forward_kernel<T, GRID_ROWS, R, G, 0, false>,
                                                      fft_forward_kernel<T, GRID_ROWS, R, G, 1, false>,
                                                      fft_forward_kernel<T, GRID_ROWS, R, G, lean_mf<T>(), false>,
                                                      fft_forward_kernel<T, GRID_ROWS, R, G, wide_mf<T>(), false>);
    if constexpr (T >= 8) {
        if (use_vec<T, R, G>(L, pm, spec, ss)) {
            k = pick_variant(f, fft_forward_kernel<T, GRID_COLS, R, G, 0, true>,
                             fft_forward_kernel<T, GRID_COLS, R, G, 1, true>,
                             fft_forward_kernel<T, GRID_COLS, R, G, lean_mf<T>(), true>,
                             fft_forward_kernel<T, GRID_COLS, R, G, wide_mf<T>(), true>);
        }
    }
    if (lds > 65536)
        check_hip(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                  "lds attribute");
    const int nt = !f.mfma_pass ? fft_threads<T, R, 0>()
                   : f.wide   ? fft_threads<T, R, wide_mf<T>()>()
                   : f.lean   ? fft_threads<T, R, lean_mf<T>()>() : fft_threads<T, R, 1>();
    const FftFwdArgs<R, G> a{f.desc, tw_of<R>(f), twp_of<R>(f), f.perm, grid, (C*)spec, ss, pm,
                             L.nm_used, L.scale, L.weighted ? L.t->wlat : nullptr, L.t->wrad,
                             L.center_factor};
    k<<<grd, nt, lds, s>>>(a);
    check_hip(hipGetLastError(), "fft_forward launch");
}

template <int T, class R, class G>
void launch_backward(const FftLaunch& L, const void* spec, long ss, Inst<G> grid, int nb,
                     hipStream_t s) {
    using C = typename Cplx<R>::type;
    const FftPlan& f = L.t->fft;
    const PairMap pm = pair_map(*L.g, T);
    const size_t lds = lds_bytes(f.desc.n, f.desc.npass, T, sizeof(C));
    dim3 grd(grid_tiles(pm), nb);
    auto k = L.layout_kind == GRID_COLS ? pick_variant(f, fft_backward_kernel<T, GRID_COLS, R, G, 0, false>,
                                                      fft_backward_kernel<T, GRID_COLS, R, G, 1, false>,
                                                      fft_backward_kernel<T, GRID_COLS, R, G, lean_mf<T>(), false>,
                                                      fft_backward_kernel<T, GRID_COLS, R, G, wide_mf<T>(), false>)
                                        : pick_variant(f, fft_backward_kernel<T, GRID_ROWS, R, G, 0, false>,
                                                      fft_backward_kernel<T, GRID_ROWS, R, G, 1, false>,
                                                      fft_backward_kernel<T, GRID_ROWS, R, G, lean_mf<T>(), false>,
                                                      fft_backward_kernel<T, GRID_ROWS, R, G, wide_mf<T>(), false>);
    if constexpr (T >= 8) {
        if (use_vec<T, R, G>(L, pm, spec, ss))
            k = pick_variant(f, fft_backward_kernel<T, GRID_COLS, R, G, 0, true>,
                             fft_backward_kernel<T, GRID_COLS, R, G, 1, true>,
                             fft_backward_kernel<T, GRID_COLS, R, G, lean_mf<T>(), true>,
                             fft_backward_kernel<T, GRID_COLS, R, G, wide_mf<T>(), true>);
    }
    if (lds > 65536)
        check_hip(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                  "lds attribute");
    const int nt = !f.mfma_pass ? fft_threads<T, R, 0>()
                   : f.wide   ? fft_threads<T, R, wide_mf<T>()>()
                   : f.lean   ? fft_threads<T, R, lean_mf<T>()>() : fft_threads<T, R, 1>();
    const FftBwdArgs<R, G> a{f.desc, tw_of<R>(f), twp_of<R>(f), f.perm, (const C*)spec, ss, grid,
                             pm, L.nm_used, L.scale, L.weighted ? L.t->wlat : nullptr, L.t->wrad,
                             L.center_factor};
    k<<<grd, nt, lds, s>>>(a);
    check_hip(hipGetLastError(), "fft_backward launch");
}

template <class R, class G>
void dispatch_forward(const FftLaunch& L, Inst<const G> grid, void* spec, long ss, int nb,
                      hipStream_t s) {
    switch (L.tile ? L.tile : L.t->fft.tile[sizeof(R) == 4]) {
        case 32: launch_forward<32, R>(L, grid, spec, ss, nb, s); break;
        case 16: launch_forward<16, R>(L, grid, spec, ss, nb, s); break;
        case 8: launch_forward<8, R>(L, grid, spec, ss, nb, s); break;
        case 4: launch_forward<4, R>(L, grid, spec, ss, nb, s); break;
        case 2: launch_forward<2, R>(L, grid, spec, ss, nb, s); break;
        default: launch_forward<1, R>(L, grid, spec, ss, nb, s); break;
    }
}

template <class R, class G>
void dispatch_backward(const FftLaunch& L, const void* spec, long ss, Inst<G> grid, int nb,
                       hipStream_t s) {
    switch (L.tile ? L.tile : L.t->fft.tile[sizeof(R) == 4]) {
        case 32: launch_backward<32, R>(L, spec, ss, grid, nb, s); break;
        case 16: launch_backward<16, R>(L, spec, ss, grid, nb, s); break;
        case 8: launch_backward<8, R>(L, spec, ss, grid, nb, s); break;
        case 4: launch_backward<4, R>(L, spec, ss, grid, nb, s); break;
        case 2: launch_backward<2, R>(L, spec, ss, grid, nb, s); break;
        default: launch_backward<1, R>(L, spec, ss, grid, nb, s); break;
    }
}

// ---- persistent launches over many orders (ragged batches) ----
// The kernel of one (width T, variant, vector form): variant 0 plain, 1 MFMA
// pass, 2 lean, 3 wide, as pick_variant chooses them.
template <int T, class R, bool VEC>
struct FftPersistentKernels {
    using FA = FftFwdArgs<R, double>;
    using BA = FftBwdArgs<R, double>;
    using FK = void (*)(const FA*, const int4*, const int*);
    using BK = void (*)(const BA*, const int4*, const int*);
    static FK fwd(int var) {
        switch (var) {
            case 0: return fft_forward_persistent<T, GRID_COLS, R, double, 0, VEC>;
            case 1: return fft_forward_persistent<T, GRID_COLS, R, double, 1, VEC>;
            case 2: return fft_forward_persistent<T, GRID_COLS, R, double, lean_mf<T>(), VEC>;
            default: return fft_forward_persistent<T, GRID_COLS, R, double, wide_mf<T>(), VEC>;
        }
    }
    static BK bwd(int var) {
        switch (var) {
            case 0: return fft_backward_persistent<T, GRID_COLS, R, double, 0, VEC>;
            case 1: return fft_backward_persistent<T, GRID_COLS, R, double, 1, VEC>;
            case 2: return fft_backward_persistent<T, GRID_COLS, R, double, lean_mf<T>(), VEC>;
            default: return fft_backward_persistent<T, GRID_COLS, R, double, wide_mf<T>(), VEC>;
        }
    }
    static int threads(int var) {
        switch (var) {
            case 0: return fft_threads<T, R, 0>();
            case 1: return fft_threads<T, R, 1>();
            case 2: return fft_threads<T, R, lean_mf<T>()>();
            default: return fft_threads<T, R, wide_mf<T>()>();
        }
    }
};

inline int fft_variant(const FftPlan& f) { return !f.mfma_pass ? 0 : f.wide ? 3 : f.lean ? 2 : 1; }

// One persistent launch over the groups `ids` (all of width T, variant var,
// vector form VEC): items (group, tile, instance), modelled cost n * T each.
template <int T, class R, bool VEC>
void fft_persistent_bucket(std::vector<PersistentLaunch>& out, const std::vector<FftGroup>& groups,
                           const std::vector<int>& ids, int var, bool forward) {
    using C = typename Cplx<R>::type;
    using K = FftPersistentKernels<T, R, VEC>;
    PersistentLaunch pl;
    pl.stage = forward ? "fft_forward" : "fft_backward";
    std::vector<typename K::FA> fa;
    std::vector<typename K::BA> ba;
    std::vector<int4> items;
    std::vector<double> cost;
    size_t lds = 0;
    for (int j = 0; j < (int)ids.size(); ++j) {
        const FftGroup& gr = groups[ids[j]];
        const FftLaunch& L = gr.L;
        const FftPlan& f = L.t->fft;
        const PairMap pm = pair_map(*L.g, T);
        lds = std::max(lds, lds_bytes(f.desc.n, f.desc.npass, T, sizeof(C)));
        if (forward)
            fa.push_back(typename K::FA{f.desc, tw_of<R>(f), twp_of<R>(f), f.perm, Inst<const double>(gr.grid),
                                        (C*)gr.spec, gr.spec_stride, pm, L.nm_used, L.scale,
                                        L.weighted ? L.t->wlat : nullptr, L.t->wrad, L.center_factor});
        else
            ba.push_back(typename K::BA{f.desc, tw_of<R>(f), twp_of<R>(f), f.perm, (const C*)gr.spec,
                                        gr.spec_stride, gr.grid, pm, L.nm_used, L.scale,
                                        L.weighted ? L.t->wlat : nullptr, L.t->wrad, L.center_factor});
        pl.instances += gr.nb;
        for (int b = 0; b < gr.nb; ++b)
            for (int t = 0; t < pm.ntiles; ++t) {
                items.push_back(make_int4(j, t, b, 0));
                cost.push_back((double)f.desc.n * T);
            }
    }
    if (items.empty()) return;
    const int nt = K::threads(var);
    const void* kern = forward ? reinterpret_cast<const void*>(K::fwd(var)) : reinterpret_cast<const void*>(K::bwd(var));
    if (lds > 65536)
        check_hip(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds), "lds attribute");
    std::vector<int4> dealt;
    std::vector<int> bounds;
    lpt_deal(items, cost, resident_blocks(kern, nt, lds), dealt, bounds, pl);
    const int blocks = (int)bounds.size() - 1;
    const int4* di = upload_owned(pl, dealt);
    const int* db = upload_owned(pl, bounds);
    pl.items = (long)items.size();
    pl.blocks = blocks;
    if (forward) {
        const typename K::FA* dg = upload_owned(pl, fa);
        const typename K::FK k = K::fwd(var);
        pl.run = [=](hipStream_t s) {
            k<<<blocks, nt, lds, s>>>(dg, di, db);
            check_hip(hipGetLastError(), "persistent fft_forward launch");
        };
    } else {
        const typename K::BA* dg = upload_owned(pl, ba);
        const typename K::BK k = K::bwd(var);
        pl.run = [=](hipStream_t s) {
            k<<<blocks, nt, lds, s>>>(dg, di, db);
            check_hip(hipGetLastError(), "persistent fft_backward launch");
        };
    }
    out.push_back(std::move(pl));
}

template <int T, class R>
void fft_persistent_width(std::vector<PersistentLaunch>& out, const std::vector<FftGroup>& groups,
                          const std::vector<int>& ids, bool forward) {
    std::map<std::pair<int, bool>, std::vector<int>> buckets;
    for (int i : ids) {
        const FftGroup& gr = groups[i];
        const PairMap pm = pair_map(*gr.L.g, T);
        bool vec = false;
        if constexpr (T >= 8) vec = use_vec<T, R, double>(gr.L, pm, gr.spec, gr.spec_stride);
        buckets[{fft_variant(gr.L.t->fft), vec}].push_back(i);
    }
    for (auto& kv : buckets) {
        if constexpr (T >= 8)
            if (kv.first.second) {
                fft_persistent_bucket<T, R, true>(out, groups, kv.second, kv.first.first, forward);
                continue;
            }
        fft_persistent_bucket<T, R, false>(out, groups, kv.second, kv.first.first, forward);
    }
}

// Widths below 8 (long sequences, orders above those of ragged batches) have
// no persistent form: false, nothing appended.
template <class R>
bool fft_persistent(std::vector<PersistentLaunch>& out, const std::vector<FftGroup>& groups, bool forward) {
    std::map<int, std::vector<int>> by_width;
    for (int i = 0; i < (int)groups.size(); ++i) {
        const FftLaunch& L = groups[i].L;
        const int T = L.tile ? L.tile : L.t->fft.tile[sizeof(R) == 4];
        if (L.layout_kind != GRID_COLS || T < 8) return false;
        by_width[T].push_back(i);
    }
    for (auto& kv : by_width) {
        switch (kv.first) {
            case 32: fft_persistent_width<32, R>(out, groups, kv.second, forward); break;
            case 16: fft_persistent_width<16, R>(out, groups, kv.second, forward); break;
            default: fft_persistent_width<8, R>(out, groups, kv.second, forward); break;
        }
    }
    return true;
}

// Whether T column pairs per workgroup fit the LDS and the per-thread item
// budget of the LDS-staged large-prime pass.
bool tile_fits(int n, const int* rad, int npass, int T, size_t csize, bool mf, bool wide = false) {
    const size_t limit = (size_t)std::min(lds_limit(), 160 * 1024);
    if (lds_bytes(n, npass, T, csize) > limit) return false;
    const bool wide_t = wide && (T == 4 || T == 8 || T == 16);  // widths with wide_mf<T>() == 3
    for (int k = 0; k < npass; ++k)
        if (rad[k] > kMaxRegRadix && !(mf && mfma_radix(rad[k])) && !(wide_t && wide_radix(rad[k]))) {
            const long items = (long)(n / rad[k]) * T * ((rad[k] - 1) / 2 + 1);
            if ((items + kThreads - 1) / kThreads > kMaxBigItems) return false;
        }
    return true;
}

// The widest tile (at most 16) that leaves room for two workgroups per CU
// (80 KB of LDS each), else the widest that fits at all.
[[maybe_unused]] int choose_tile(int n, const std::vector<int>& rad, size_t csize, bool mf, bool wide) {
    const size_t limit = (size_t)std::min(lds_limit(), 160 * 1024);
    const int npass = (int)rad.size();
    for (int T = 16; T >= 1; T /= 2)
        if (lds_bytes(n, npass, T, csize) <= 80 * 1024 && tile_fits(n, rad.data(), npass, T, csize, mf, wide))
            return T;
    int T = 16;
    for (; T > 1; T /= 2)
        if (tile_fits(n, rad.data(), npass, T, csize, mf, wide)) break;
    if (lds_bytes(n, npass, T, csize) > limit)
        throw Error{ZGL_ERROR_NOT_SUPPORTED, "longitude count too large for the LDS"};
    return T;
}

template <class C>
C* upload(const std::vector<C>& v, hipStream_t s, const char* what) {
    C* d = dmalloc_n<C>(v.size(), what);
    try {
        if (!v.empty())
            check_hip(hipMemcpyAsync(d, v.data(), v.size() * sizeof(C), hipMemcpyHostToDevice, s), what);
    } catch (...) {
        dfree(d);
        throw;
    }
    return d;
}

}  // namespace

#if ZGL_FFT_PART == 0
std::vector<int> fft_tiles(const FftPlan& p, bool fp32) {
    // the plan's width (first: the autotuner's preference on a near tie),
    // the double width where it fits (longer contiguous segments per
    // longitude row, fewer workgroups per CU) and half of it
    std::vector<int> out;
    const size_t cs = fp32 ? sizeof(float2) : sizeof(double2);
    const int T0 = p.tile[fp32 ? 1 : 0];
    auto fits = [&](int T) { return tile_fits(p.desc.n, p.desc.radix, p.desc.npass, T, cs, p.mfma_pass, p.wide); };
    if (fits(T0)) out.push_back(T0);
    if (T0 < 32 && fits(2 * T0)) out.push_back(2 * T0);
    for (int T = T0 / 2; T >= 1 && out.size() < 3; T /= 2)
        if (fits(T)) out.push_back(T);
    return out;
}

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
    // prime radices in (17, kMaxMfmaRadix] run on the matrix cores, and those
    // in (kMaxMfmaRadix, kMaxWideRadix] too when every other radix is lean
    // (the 512-thread variant), unless ZGL_FFT_MFMA=0 (comparison runs)
    bool has_wide = false, lean_all = true;
    for (int r : rad) {
        p.mfma_pass |= mfma_radix(r);
        has_wide |= wide_radix(r);
        lean_all &= r <= kLeanRadix || mfma_radix(r) || wide_radix(r);
    }
    p.wide = has_wide && lean_all;
    p.mfma_pass |= p.wide;
    if (const char* e = std::getenv("ZGL_FFT_MFMA"))
        if (e[0] == '0') p.mfma_pass = p.wide = false;
    p.lean = p.mfma_pass;
    for (int r : rad) p.lean &= r <= kLeanRadix || mfma_radix(r) || (p.wide && wide_radix(r));
    const bool mf = p.mfma_pass;
    p.tile[0] = choose_tile(n, rad, sizeof(double2), mf, p.wide);
    p.tile[1] = choose_tile(n, rad, sizeof(float2), mf, p.wide);
    // ZGL_FFT_TILE=T forces a width where it fits (comparison and test runs)
    if (const char* e = std::getenv("ZGL_FFT_TILE")) {
        const int T = std::atoi(e);
        if (T == 1 || T == 2 || T == 4 || T == 8 || T == 16 || T == 32) {
            if (tile_fits(n, rad.data(), (int)rad.size(), T, sizeof(double2), mf, p.wide)) p.tile[0] = T;
            if (tile_fits(n, rad.data(), (int)rad.size(), T, sizeof(float2), mf, p.wide)) p.tile[1] = T;
        }
    }

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
    const long double two_pi = 6.283185307179586476925286766559005768L;
    auto root = [&](long t) {  // e^{-2 pi i t / n}
        const long double ang = two_pi * (long double)(t % n) / (long double)n;
        return std::pair<double, double>((double)cosl(ang), (double)(-sinl(ang)));
    };
    std::vector<double2> tw(n);
    std::vector<float2> tw32(n);
    for (int t = 0; t < n; ++t) {
        auto [c, sn] = root(t);
        tw[t] = make_double2(c, sn);
        tw32[t] = make_float2((float)c, (float)sn);
    }
    // Pass ps >= 1 of radix P over sub-transforms of length Lp: slot
    // pos = B P Lp + r Lp + q takes e^{-2 pi i r q / (P Lp)}.
    std::vector<double2> twp;
    std::vector<float2> twp32;
    for (int ps = 1; ps < p.desc.npass; ++ps) {
        const int P = rad[ps], Lp = p.desc.L[ps];
        for (int pos = 0; pos < n; ++pos) {
            const int q = pos % Lp, r = (pos / Lp) % P;
            auto [c, sn] = root((long)r * q * (n / (P * Lp)));
            twp.push_back(make_double2(c, sn));
            twp32.push_back(make_float2((float)c, (float)sn));
        }
    }
    p.tw = upload(tw, s, "fft twiddles");
    p.tw32 = upload(tw32, s, "fft twiddles");
    p.twp = upload(twp, s, "fft pass twiddles");
    p.twp32 = upload(twp32, s, "fft pass twiddles");
    p.perm = upload(perm, s, "fft permutation");
    check_hip(hipStreamSynchronize(s), "fft plan");
}

void fft_plan_destroy(FftPlan& p) {
    for (void* q : {(void*)p.tw, (void*)p.tw32, (void*)p.twp, (void*)p.twp32, (void*)p.perm}) dfree(q);
    p = FftPlan{};
}

#endif  // ZGL_FFT_PART == 0

template <class R, class G>
void fft_forward(const FftLaunch& L, Inst<const G> grid, void* spec, long spec_stride, int nb,
                 hipStream_t s) {
    dispatch_forward<R>(L, grid, spec, spec_stride, nb, s);
}

template <class R, class G>
void fft_backward(const FftLaunch& L, const void* spec, long spec_stride, Inst<G> grid, int nb,
                  hipStream_t s) {
    dispatch_backward<R>(L, spec, spec_stride, grid, nb, s);
}

template <class R>
bool fft_forward_persistent(std::vector<PersistentLaunch>& out, const std::vector<FftGroup>& g) {
    return fft_persistent<R>(out, g, true);
}

template <class R>
bool fft_backward_persistent(std::vector<PersistentLaunch>& out, const std::vector<FftGroup>& g) {
    return fft_persistent<R>(out, g, false);
}

// FP64 grids with FP64 or FP32 arithmetic; FP32 grids with FP32 arithmetic.
#if ZGL_FFT_PART == 0
template bool fft_forward_persistent<double>(std::vector<PersistentLaunch>&, const std::vector<FftGroup>&);
template bool fft_backward_persistent<double>(std::vector<PersistentLaunch>&, const std::vector<FftGroup>&);
template void fft_forward<double, double>(const FftLaunch&, Inst<const double>, void*, long, int, hipStream_t);
template void fft_backward<double, double>(const FftLaunch&, const void*, long, Inst<double>, int, hipStream_t);
#elif ZGL_FFT_PART == 1
template bool fft_forward_persistent<float>(std::vector<PersistentLaunch>&, const std::vector<FftGroup>&);
template bool fft_backward_persistent<float>(std::vector<PersistentLaunch>&, const std::vector<FftGroup>&);
template void fft_forward<float, double>(const FftLaunch&, Inst<const double>, void*, long, int, hipStream_t);
template void fft_backward<float, double>(const FftLaunch&, const void*, long, Inst<double>, int, hipStream_t);
#else
template void fft_forward<float, float>(const FftLaunch&, Inst<const float>, void*, long, int, hipStream_t);
template void fft_backward<float, float>(const FftLaunch&, const void*, long, Inst<float>, int, hipStream_t);
#endif

}  // namespace zgl
