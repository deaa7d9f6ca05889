// Streaming MFMA kernels for the Zernike colatitude (Legendre) stages.
//
// Replaces the hot loops
//   integrate_latitudinal  include/zernike_glq_transformer.hpp:778-817
//   sum_l                  include/zernike_glq_transformer.hpp:901-941
//
// Each (m, parity) problem multiplies a small table block A (the P_lm values,
// at most 16*NS rows by jh half latitudes) with a long, contiguous operand B
// (the spectrum of one instance: jh rows of 2*ncol reals, or the flm rows of
// one instance). The roles are fixed by HBM traffic: A is staged once per
// workgroup in LDS, B and the output stream through registers only.
//
// Column permutation. An MFMA 16x16x4 takes one B value per lane, lane l
// holding B[k = l/16][j = l%16]. Any permutation of the output columns is
// allowed as long as the same permutation is used when storing, so a lane
// loads VW consecutive columns with one 16-byte load,
//   B[k][c0 + VW*j + v], v = 0..VW-1,
// and feeds value v to MFMA "virtual block" v. One load instruction then
// covers 4 rows x 256 contiguous bytes, and the results come back as VW
// consecutive columns per lane, stored with one 16-byte store. Lanes past the
// end of the column axis (a partial last tile) load zeros and store nothing.
// The number of active 16-row strips is a template parameter of the tile
// loop, dispatched once per workgroup, so the MFMA sequence is branch-free.
//
// Pipelining: each wave walks its tiles in chunks of KC k-steps, holding two
// register buffers of B values; the loads of the next chunk are in flight
// while the MFMAs of the current one run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "zgl_internal.hpp"
#include "zgl_plm.hpp"

namespace zgl {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
typedef double d4 __attribute__((ext_vector_type(4)));

template <class T> struct SM;
template <> struct SM<float> {
    using V = f4;
    using acc_t = f4;
    static constexpr int VW = 4;
    __device__ static acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // lane holds D[4(lane/16) + r][lane%16]
    __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <> struct SM<double> {
    using V = d2;
    using acc_t = d4;
    static constexpr int VW = 2;
    __device__ static acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // lane holds D[lane/16 + 4r][lane%16]
    __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

constexpr int kThreads = 256;
constexpr int kKC = 6;  // k-steps of 4 per register chunk (K = 65 -> 3 chunks)
constexpr int kMaxGroup = 16;  // instances per workgroup

// Raw buffer accesses (a wave-uniform base and byte range, 32-bit lane
// offsets): a lane whose offset lies past the range loads zeros and its store
// is dropped by the hardware, so masked accesses need no branch and every
// lane issues the same loads and stores -- the counts of memory operations in
// flight stay static, and a wait for a load never waits for stores issued
// after it.
typedef unsigned int bu4 __attribute__((vector_size(16)));
constexpr unsigned kOutOfRange = 0x80000000u;  // past every range used here (< 2 GiB)
__device__ inline const void* wave_uniform(const void* p) {
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return reinterpret_cast<const void*>(((unsigned long long)hi << 32) | lo);
}
__device__ inline __amdgpu_buffer_rsrc_t buffer_range(const void* base, unsigned bytes) {
    // dword 3 of a raw gfx9 buffer resource: 32-bit data format, no swizzle
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(wave_uniform(base)), 0, (int)bytes, 0x00020000);
}
template <class V>
__device__ inline V buffer_load16(__amdgpu_buffer_rsrc_t r, unsigned off) {
    static_assert(sizeof(V) == 16, "16-byte accesses");
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}
template <class V>
__device__ inline void buffer_store16(__amdgpu_buffer_rsrc_t r, unsigned off, V v) {
    static_assert(sizeof(V) == 16, "16-byte accesses");
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(bu4, v), r, (int)off, 0, 0);
}
typedef unsigned int bu2 __attribute__((vector_size(8)));
// one (re, im) pair of T: 8 bytes (float) or 16 bytes (double)
template <class T>
__device__ inline void buffer_store_pair(__amdgpu_buffer_rsrc_t r, unsigned off, T re, T im) {
    if constexpr (sizeof(T) == 8) {
        typedef double dd2 __attribute__((ext_vector_type(2)));
        const dd2 v = {re, im};
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(bu4, v), r, (int)off, 0, 0);
    } else {
        typedef float ff2 __attribute__((ext_vector_type(2)));
        const ff2 v = {re, im};
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(bu2, v), r, (int)off, 0, 0);
    }
}

// Fills n elements of a workgroup's table block in LDS from global memory:
// map(e) gives element e's LDS index and source (ok false: a zero, src then
// any readable address); consecutive e should be consecutive in memory. B loads per thread are issued before any of them
// is used, so the fill costs a few memory latencies instead of one per
// element (a load under a branch inside the loop would serialise them).
template <class T>
struct FillSrc {
    const T* src;
    bool ok;
    int dst;  // LDS index
};
template <int B, class T, class F>
__device__ __forceinline__ void lds_fill(T* lds, int n, F&& map) {
    for (int e0 = threadIdx.x; e0 < n; e0 += B * kThreads) {
        T v[B];
        bool ok[B];
        int d[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const FillSrc<T> f = map(min(e0 + u * kThreads, n - 1));
            ok[u] = f.ok;
            d[u] = f.dst;
            v[u] = *f.src;
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int e = e0 + u * kThreads;
            if (e < n) lds[d[u]] = ok[u] ? v[u] : T(0);
        }
    }
}

// Legendre forward: flm[l,m][c] = sum_h P_lm(z_h) X^par_m[h][c], l = m+par+2i.
template <class TT>
struct LegFwd {
    using T = TT;
    const T* ptab; const T* spec; T* flm;
    long spec_stride, flm_stride;  // elements of T
    int order, nt, jh, ncol, G, as;
    const double* zh; int qm; double phase;  // for P blocks recomputed in the kernel (ptab null)
    static constexpr bool kCenter = false;
    static constexpr bool kAccum = false;
    __device__ int center_lds_row() const { return 16 * kNsLegendreFwd; }
    struct Prob {
        int M, K, N, m, par; const T* A; long a_si, a_sk; long boff;
        bool center; const T* Ac; int hoff;
    };
    __device__ Prob problem(int prob, int m0) const {
        Prob p;
        p.m = prob >> 1; p.par = prob & 1;
        p.M = plm_count(nt, p.m, p.par); p.K = jh; p.N = 2 * ncol;
        p.center = false; p.Ac = nullptr; p.hoff = 0;
        p.A = ptab ? ptab + plm_offset(order, p.m, p.par, jh) + (long)m0 * jh : nullptr;
        p.a_si = jh; p.a_sk = 1;
        p.boff = (long)(2 * p.m + p.par) * jh * 2 * ncol;
        return p;
    }
    // operand rows as a per-instance base (wave-uniform) plus a 32-bit offset
    __device__ const T* b_base(int inst) const { return spec + inst * spec_stride; }
    __device__ unsigned b_off(const Prob& p, int k) const { return (unsigned)(p.boff + k * 2 * ncol); }
    __device__ T* c_base(int inst) const { return flm + inst * flm_stride; }
    __host__ __device__ long b_stride() const { return spec_stride; }
    __host__ __device__ long c_stride() const { return flm_stride; }
    __device__ unsigned c_off(const Prob& p, int i) const {
        return (unsigned)(tri_idx(p.m + p.par + 2 * i, p.m) * 2 * ncol);
    }
    // A block by recursion (zeroed beforehand): As[i][h] = P_l(z_h),
    // l = m + par + 2 (m0 + i), one thread per half latitude h
    __device__ void plm_fill(T* As, int AS, const Prob& p, int m0, int rows, int TM) const {
        for (int h = threadIdx.x; h < p.K; h += blockDim.x)
            plm_column(order, p.m, zh[h], qm, phase, [&](int l, double v) {
                const int d = l - p.m - p.par;
                if (d >= 0 && !(d & 1)) {
                    const int i = (d >> 1) - m0;
                    if (i >= 0 && i < rows) As[i * AS + h] = (T)v;
                }
            });
    }
};

// Legendre backward: X^par_m[h][c] = sum_i P_lm(z_h) flm[l,m][c], l = m+par+2i.
// Half latitude h pairs latitudes nth/2 + h and jh - 1 - h; with odd nth,
// h = 0 is the equator, paired with itself. The MFMA rows are the nth/2 half
// latitudes off the equator, h = hoff + r with hoff = nth & 1; the equator row
// is computed beside them on the VALU by the first row block of parity 0.
// Parity 1 has no equator term (the backward FFT reads only the even part
// there).
template <class TT>
struct LegBwd {
    using T = TT;
    const T* ptab; const T* flm; T* spec;
    long spec_stride, flm_stride;
    int order, nt, jh, ncol, G, as, nth;
    const double* zh; int qm; double phase;  // for P blocks recomputed in the kernel (ptab null)
    static constexpr bool kCenter = true;
    static constexpr bool kAccum = false;
    __device__ int center_lds_row() const { return 16 * kNsStreamBwd; }
    struct Prob {
        int M, K, N, m, par; const T* A; long a_si, a_sk; long boff;
        bool center; const T* Ac; int hoff;
    };
    __device__ Prob problem(int prob, int m0) const {
        Prob p;
        p.m = prob >> 1; p.par = prob & 1;
        p.M = p.m < nt ? nth / 2 : 0;
        p.K = plm_count(nt, p.m, p.par); p.N = 2 * ncol;
        p.hoff = nth & 1;
        const T* a0 = ptab ? ptab + plm_offset(order, p.m, p.par, jh) : nullptr;  // [i][h]
        p.A = a0 ? a0 + p.hoff + m0 : nullptr;
        p.a_si = 1; p.a_sk = jh;
        p.boff = (long)(2 * p.m + p.par) * jh * 2 * ncol;
        p.center = p.hoff && p.par == 0 && m0 == 0;
        p.Ac = a0;  // the equator column h = 0
        return p;
    }
    __device__ const T* b_base(int inst) const { return flm + inst * flm_stride; }
    __device__ unsigned b_off(const Prob& p, int k) const {
        return (unsigned)(tri_idx(p.m + p.par + 2 * k, p.m) * 2 * ncol);
    }
    __device__ T* c_base(int inst) const { return spec + inst * spec_stride; }
    __host__ __device__ long b_stride() const { return flm_stride; }
    __host__ __device__ long c_stride() const { return spec_stride; }
    // output row r is half latitude h = r + hoff (r = -hoff: the equator)
    __device__ unsigned c_off(const Prob& p, int r) const {
        return (unsigned)(p.boff + (r + p.hoff) * 2 * ncol);
    }
    // A block by recursion (zeroed beforehand): As[r][k] = P_l(z_h), h = hoff +
    // m0 + r, l = m + par + 2k; the equator row (P.center) goes to As row TM
    __device__ void plm_fill(T* As, int AS, const Prob& p, int m0, int rows, int TM) const {
        const int nrow = rows + (p.center ? 1 : 0);
        for (int r = threadIdx.x; r < nrow; r += blockDim.x) {
            const int h = r < rows ? p.hoff + m0 + r : 0;
            T* dst = As + (long)(r < rows ? r : TM) * AS;
            plm_column(order, p.m, zh[h], qm, phase, [&](int l, double v) {
                const int d = l - p.m - p.par;
                if (d >= 0 && !(d & 1) && (d >> 1) < p.K) dst[d >> 1] = (T)v;
            });
        }
    }
};

// Evaluation at arbitrary colatitudes (zt::GridEvaluator::sum_l,
// src/grid_evaluator.cpp:185-219; st::GridEvaluator::sum_l,
// include/grid_evaluator.hpp:136-159):
//   S[m][j][c] = sum_{l=m}^{nt-1} P_lm(z_j) flm[l,m][c].
// The points have no north/south symmetry, so one problem per m takes every
// l (no parity split) and the rows are the nlat points. P comes from the
// table [m][l-m][nlat] or is recomputed from z in the workgroup.
template <class TT>
struct LegEval {
    using T = TT;
    const T* ptab; const T* flm; T* spec;
    long spec_stride, flm_stride;
    int order, nt, jh, ncol, G, as;  // jh: the number of points nlat
    const double* zh; int qm; double phase;
    int l0, kchunk;  // this pass sums l = m + l0 .. m + l0 + kchunk - 1
    int accum;       // add to S (passes after the first)
    static constexpr bool kCenter = false;
    static constexpr bool kAccum = true;
    __device__ int center_lds_row() const { return 0; }
    struct Prob {
        int M, K, N, m, par; const T* A; long a_si, a_sk; long boff;
        bool center; const T* Ac; int hoff;
    };
    __device__ Prob problem(int prob, int m0) const {
        Prob p;
        p.m = prob; p.par = 0;
        p.K = min(kchunk, nt - p.m - l0);
        p.M = p.K > 0 ? jh : 0;
        p.N = 2 * ncol;
        p.hoff = 0; p.center = false; p.Ac = nullptr;
        // rows of m' < m: sum (order - m') of them, each jh long
        const long off = ((long)p.m * order - ((long)p.m * (p.m - 1)) / 2 + l0) * jh;
        p.A = ptab ? ptab + off + m0 : nullptr;
        p.a_si = 1; p.a_sk = jh;
        p.boff = (long)p.m * jh * 2 * ncol;
        return p;
    }
    __device__ const T* b_base(int inst) const { return flm + inst * flm_stride; }
    __device__ unsigned b_off(const Prob& p, int k) const {
        return (unsigned)(tri_idx(p.m + l0 + k, p.m) * 2 * ncol);
    }
    __device__ T* c_base(int inst) const { return spec + inst * spec_stride; }
    __host__ __device__ long b_stride() const { return flm_stride; }
    __host__ __device__ long c_stride() const { return spec_stride; }
    __device__ unsigned c_off(const Prob& p, int r) const { return (unsigned)(p.boff + (long)r * 2 * ncol); }
    // As[r][k] = P_{m+l0+k,m}(z_{m0+r})
    __device__ void plm_fill(T* As, int AS, const Prob& p, int m0, int rows, int) const {
        for (int r = threadIdx.x; r < rows; r += blockDim.x) {
            T* dst = As + (long)r * AS;
            plm_column(order, p.m, zh[m0 + r], qm, phase, [&](int l, double v) {
                const int k = l - p.m - l0;
                if (k >= 0 && k < p.K) dst[k] = (T)v;
            });
        }
    }
};

// The tile loop of one wave, for NSA active 16-row strips (a compile-time
// count, so the MFMA sequence has no per-strip branches). The workgroup's
// tiles start at column c0 of instance g0 and run over the flattened
// (instance, column) index.
template <int NSA, bool CENTER, class S>
__device__ __forceinline__ void stream_tiles(const S& st, const typename S::Prob& P,
                                             const typename S::T* As, int AS, int m0, int rows,
                                             int g0, int gn, int c0, int tiles, int wave, int lane) {
    using T = typename S::T;
    using M = SM<T>;
    using V = typename M::V;
    constexpr int VW = M::VW;
    constexpr int TNW = 16 * VW;
    const int j = lane & 15, g = lane >> 4;
    const int K = P.K, steps = (K + 3) >> 2;
    const int N = P.N;
    const int ntw = tiles > wave ? (tiles - wave + 3) >> 2 : 0;  // tiles of this wave
    const int nchunks = max(1, (steps + kKC - 1) / kKC);
    const int U = ntw * nchunks;

    typename M::acc_t acc[NSA][VW];
#pragma unroll
    for (int q = 0; q < NSA; ++q)
#pragma unroll
        for (int v = 0; v < VW; ++v) acc[q][v] = typename M::acc_t{0, 0, 0, 0};
    // equator row (CENTER): per-lane partial sums over this lane's k values;
    // its A row is As row 16 * NS of the kernel
    V cpart = {};
    const T* Ac = As + (long)st.center_lds_row() * AS;

    // unit u -> (instance, first column of this lane, chunk). A tile's TNW
    // columns run over the flattened (instance, column) index of the group,
    // so a tile may end one instance and start the next and only the
    // workgroup's last tile can be partial (N and c0 are multiples of 16, so a lane's
    // VW columns never straddle two instances); lanes past the group get
    // inst >= g0 + gn
    auto unit = [&](int u, int& inst, int& col, int& ch) {
        const int tw = u / nchunks;
        ch = u - tw * nchunks;
        const int fc = c0 + (wave + 4 * tw) * TNW + VW * j;
        const int gi = fc / N;
        inst = g0 + gi;
        col = fc - gi * N;
    };

    // Buffer accesses relative to the group's first instance (the host keeps
    // a group's operand and output spans below 2 GiB): lanes past the group
    // or the K axis load zeros, so every lane issues all kKC loads, the
    // number of loads in flight is static, and the MFMAs of a chunk wait
    // only for their own operands.
    const __amdgpu_buffer_rsrc_t rb =
        buffer_range(st.b_base(g0), (unsigned)(gn * st.b_stride() * (long)sizeof(T)));
    const __amdgpu_buffer_rsrc_t rc =
        buffer_range(st.c_base(g0), (unsigned)(gn * st.c_stride() * (long)sizeof(T)));
    auto load = [&](V (&buf)[kKC], int u) {
        int inst, col, ch;
        unit(u, inst, col, ch);
        const bool colok = inst < g0 + gn;
        const unsigned ib = (unsigned)((inst - g0) * st.b_stride()) + col;
#pragma unroll
        for (int s = 0; s < kKC; ++s) {
            const int k = 4 * (ch * kKC + s) + g;
            const bool ok = colok && k < K;
            buf[s] = buffer_load16<V>(rb, ok ? (ib + st.b_off(P, k)) * (unsigned)sizeof(T) : kOutOfRange);
        }
    };

    // Every chunk issues its tile's stores -- dropped (out of range) except
    // after the tile's last chunk -- so the count of stores between a
    // prefetch and its use is static and its wait never covers stores.
    auto compute = [&](const V (&buf)[kKC], int u) {
        int inst, col, ch;
        unit(u, inst, col, ch);
        const bool colok = inst < g0 + gn;
#pragma unroll
        for (int s = 0; s < kKC; ++s) {
            const int ks = ch * kKC + s;
            if (ks < steps) {
#pragma unroll
                for (int q = 0; q < NSA; ++q) {
                    const T a = As[(16 * q + j) * AS + 4 * ks + g];
#pragma unroll
                    for (int v = 0; v < VW; ++v) acc[q][v] = M::mma(a, buf[s][v], acc[q][v]);
                }
                if constexpr (CENTER) {
                    const T a = Ac[4 * ks + g];
#pragma unroll
                    for (int v = 0; v < VW; ++v) cpart[v] += a * buf[s][v];
                }
            }
        }
        const bool last = ch == nchunks - 1;
        if constexpr (S::kAccum) {
            // evaluation passes after the first add to the stored sums
            if (!last) return;
            T* cb = st.c_base(inst);
#pragma unroll
            for (int q = 0; q < NSA; ++q) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = 16 * q + M::row(lane, r);
                    if (colok && i < rows) {
                        V o;
#pragma unroll
                        for (int v = 0; v < VW; ++v) o[v] = acc[q][v][r];
                        V* dst = reinterpret_cast<V*>(cb + (st.c_off(P, m0 + i) + col));
                        if (st.accum) o += *dst;
                        *dst = o;
                    }
                }
#pragma unroll
                for (int v = 0; v < VW; ++v) acc[q][v] = typename M::acc_t{0, 0, 0, 0};
            }
        } else {
            const unsigned ic = (unsigned)((inst - g0) * st.c_stride()) + col;
#pragma unroll
            for (int q = 0; q < NSA; ++q) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = 16 * q + M::row(lane, r);
                    V o;
#pragma unroll
                    for (int v = 0; v < VW; ++v) o[v] = acc[q][v][r];
                    const bool ok = last && colok && i < rows;
                    buffer_store16(rc, ok ? (ic + st.c_off(P, m0 + i)) * (unsigned)sizeof(T) : kOutOfRange, o);
                }
                if (last)
#pragma unroll
                    for (int v = 0; v < VW; ++v) acc[q][v] = typename M::acc_t{0, 0, 0, 0};
            }
        }
        if constexpr (CENTER) {
            // sum the four lane groups (k = 4s + g) and store from group 0
            V c = cpart;
#pragma unroll
            for (int v = 0; v < VW; ++v) {
                T x = c[v];
                x += __shfl_xor(x, 16);
                x += __shfl_xor(x, 32);
                c[v] = x;
            }
            const bool ok = last && colok && g == 0;
            buffer_store16(rc, ok ? ((unsigned)((inst - g0) * st.c_stride()) + col + st.c_off(P, -P.hoff)) *
                                        (unsigned)sizeof(T)
                                  : kOutOfRange,
                           c);
            if (last) cpart = V{};
        }
    };

    // the prefetch is unconditional too: past the last unit it re-reads that
    // unit (cache hits) and the values are not used
    if (U <= 0) return;
    V b0[kKC], b1[kKC];
    load(b0, 0);
    for (int u = 0; u < U; u += 2) {
        load(b1, min(u + 1, U - 1));
        compute(b0, u);
        if (u + 1 >= U) break;
        load(b0, min(u + 2, U - 1));
        compute(b1, u + 1);
    }
}

// ---------------------------------------------------------------- BF16x3
// The FP32 tile loop with the products formed on the BF16 MFMA: each FP32
// operand is split as x = hi + mid + lo (three BF16 values, round to nearest
// even, ~24 bits together) and a*b is accumulated in FP32 from the six
// leading partial products (lo*hi, mid*mid, hi*lo, mid*hi, hi*mid, hi*hi,
// smallest first), dropping terms below 2^-24 relative. MFMA
// v_mfma_f32_16x16x16_bf16: lane l holds A[l % 16][4 (l / 16) + i] and
// B[4 (l / 16) + i][l % 16], i < 4, and D like the FP32 16x16x4 MFMA, so a
// lane streams the same 16-byte column loads for four consecutive k (a step
// covers 16 rows) and stores exactly as the FP32 loop does.
typedef short s4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short bf16_rne(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf16_value(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
__device__ inline void split3(float x, short& h, short& m, short& l) {
    const unsigned short a = bf16_rne(x);
    const float r1 = x - bf16_value(a);
    const unsigned short b = bf16_rne(r1);
    const float r2 = r1 - bf16_value(b);
    h = (short)a;
    m = (short)b;
    l = (short)bf16_rne(r2);
}

constexpr int kKC3 = 1;  // 16-row steps per register chunk

template <int NSA, bool CENTER, class S>
__device__ __forceinline__ void stream_tiles_x3(const S& st, const typename S::Prob& P, const float* As, int AS,
                                                int m0, int rows, int g0, int gn, int wave, int lane) {
    using M = SM<float>;
    using V = f4;
    constexpr int VW = 4;
    constexpr int TNW = 16 * VW;
    const int j = lane & 15, g = lane >> 4;
    const int K = P.K, steps = (K + 15) >> 4;
    const int N = P.N;
    const int ntile = (N + TNW - 1) / TNW;
    const int tiles = gn * ntile;
    const int ntw = tiles > wave ? (tiles - wave + 3) >> 2 : 0;
    const int nchunks = max(1, (steps + kKC3 - 1) / kKC3);
    const int U = ntw * nchunks;

    f4 acc[NSA][VW];
#pragma unroll
    for (int q = 0; q < NSA; ++q)
#pragma unroll
        for (int v = 0; v < VW; ++v) acc[q][v] = f4{0, 0, 0, 0};
    V cpart = {};
    const float* Ac = As + (long)st.center_lds_row() * AS;

    auto unit = [&](int u, int& inst, int& col, int& ch) {
        const int tw = u / nchunks;
        ch = u - tw * nchunks;
        const int t = wave + 4 * tw;
        const int gi = t / ntile;
        inst = g0 + gi;
        col = (t - gi * ntile) * TNW + VW * j;
    };
    // rows k = 16 (chunk step) + 4 g + r, r < 4
    // branch-free loads, as in stream_tiles
    auto load = [&](V (&buf)[kKC3][4], int u) {
        int inst, col, ch;
        unit(u, inst, col, ch);
        const bool colok = col < N;
        const float* bb = st.b_base(inst);
#pragma unroll
        for (int s = 0; s < kKC3; ++s)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 16 * (ch * kKC3 + s) + 4 * g + r;
                const bool ok = colok && k < K;
                buf[s][r] = *reinterpret_cast<const V*>(bb + (ok ? st.b_off(P, k) + col : 0u));
            }
    };
    auto compute = [&](const V (&buf)[kKC3][4], int u) {
        int inst, col, ch;
        unit(u, inst, col, ch);
#pragma unroll
        for (int s = 0; s < kKC3; ++s) {
            const int ks = ch * kKC3 + s;
            if (ks < steps) {
                V b[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) b[r] = (col < N && 16 * ks + 4 * g + r < K) ? buf[s][r] : V{};
                s4 bh[VW], bm[VW], bl[VW];
#pragma unroll
                for (int v = 0; v < VW; ++v)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        short h, m, l;
                        split3(b[r][v], h, m, l);
                        bh[v][r] = h;
                        bm[v][r] = m;
                        bl[v][r] = l;
                    }
#pragma unroll
                for (int q = 0; q < NSA; ++q) {
                    s4 ah, am, al;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        short h, m, l;
                        split3(As[(16 * q + j) * AS + 16 * ks + 4 * g + r], h, m, l);
                        ah[r] = h;
                        am[r] = m;
                        al[r] = l;
                    }
#pragma unroll
                    for (int v = 0; v < VW; ++v) {
                        f4 c = acc[q][v];
                        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh[v], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am, bm[v], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl[v], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am, bh[v], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bm[v], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh[v], c, 0, 0, 0);
                        acc[q][v] = c;
                    }
                }
                if constexpr (CENTER) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float a = Ac[16 * ks + 4 * g + r];
#pragma unroll
                        for (int v = 0; v < VW; ++v) cpart[v] += a * b[r][v];
                    }
                }
            }
        }
        if (ch != nchunks - 1) return;
        const bool colok = col < N;
        float* cb = st.c_base(inst);
#pragma unroll
        for (int q = 0; q < NSA; ++q) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * q + M::row(lane, r);
                if (colok && i < rows) {
                    V o;
#pragma unroll
                    for (int v = 0; v < VW; ++v) o[v] = acc[q][v][r];
                    V* dst = reinterpret_cast<V*>(cb + (st.c_off(P, m0 + i) + col));
                    if constexpr (S::kAccum)
                        if (st.accum) o += *dst;
                    *dst = o;
                }
            }
#pragma unroll
            for (int v = 0; v < VW; ++v) acc[q][v] = f4{0, 0, 0, 0};
        }
        if constexpr (CENTER) {
#pragma unroll
            for (int v = 0; v < VW; ++v) {
                float x = cpart[v];
                x += __shfl_xor(x, 16);
                x += __shfl_xor(x, 32);
                cpart[v] = x;
            }
            if (colok && g == 0) *reinterpret_cast<V*>(cb + (st.c_off(P, -P.hoff) + col)) = cpart;
            cpart = V{};
        }
    };

    if (U <= 0) return;
    V b0[kKC3][4], b1[kKC3][4];
    load(b0, 0);
    for (int u = 0; u < U; u += 2) {
        load(b1, min(u + 1, U - 1));
        compute(b0, u);
        if (u + 1 >= U) break;
        load(b0, min(u + 2, U - 1));
        compute(b1, u + 1);
    }
}

// PLM: the P block is recomputed in the workgroup (no table)
// X3: the BF16x3 tile loop (FP32 stages only)
// One work item wk (problem, first row) for instance group `ig` of nb
// instances: st.G instances, or with tpw > 0 the ig-th run of 4 tpw tiles of
// the flattened (instance, column) index of all nb instances -- the same
// number of tiles for every wave of the workgroup, whatever the row length.
template <class S, int NS, bool PLM, bool X3>
__device__ __forceinline__ void stream_body(const S& st, const int4 wk, const int ig, const int nb,
                                            const int tpw = 0) {
    using T = typename S::T;
    constexpr int TM = 16 * NS;
    constexpr int TNW = 16 * (X3 ? SM<float>::VW : SM<T>::VW);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* As = reinterpret_cast<T*>(smem_raw);

    const int m0 = wk.y;
    const typename S::Prob P = st.problem(wk.x, m0);
    const int rows = min(TM, P.M - m0);
    if (rows <= 0) return;
    int g0, gn, c0, tiles;
    if (tpw > 0) {
        const long f0 = (long)ig * 4 * tpw * TNW;  // first flattened column
        const long f1 = min((long)nb * P.N, f0 + 4L * tpw * TNW);
        if (f0 >= f1) return;
        g0 = (int)(f0 / P.N);
        c0 = (int)(f0 - (long)g0 * P.N);
        gn = (int)((f1 - 1) / P.N) - g0 + 1;
        tiles = (int)((f1 - f0 + TNW - 1) / TNW);
    } else {
        g0 = ig * st.G;
        gn = min(st.G, nb - g0);
        if (gn <= 0) return;
        c0 = 0;
        tiles = (gn * P.N + TNW - 1) / TNW;
    }

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int K = P.K;
    const int AS = st.as;

    // A tile: TM rows x KP (K rounded up to the step of the tile loop), zero padded, As[i * AS + k]; the equator row
    // (P.center) goes to As row TM. Without a table the block is recomputed.
    const bool center = S::kCenter && P.center;
    if constexpr (PLM) {
        const int nfill = (TM + (center ? 1 : 0)) * AS;
        for (int e = tid; e < nfill; e += kThreads) As[e] = T(0);
        __syncthreads();
        st.plm_fill(As, AS, P, m0, rows, TM);
    } else {
        // rows of AS (the padding columns past KP are never read; they get
        // zeros too); the equator row (P.center) is row TM
        const int nrow = TM + (center ? 1 : 0);
        // zeros read the table's first value: P.A itself lies past the end of
        // the table for a problem without rows (K = 0)
        // e runs along the table's contiguous axis: k for row-major blocks
        // (a_sk = 1), the rows i otherwise
        const bool kfast = P.a_sk == 1;
        lds_fill<8>(As, nrow * AS, [&](int e) {
            const int i = kfast ? e / AS : e % nrow, k = kfast ? e - i * AS : e / nrow;
            const bool eq = i == TM;
            const bool ok = k < K && (eq || i < rows);
            const T* src = (eq ? P.Ac : P.A) + (eq ? 0L : (long)i * P.a_si) + (long)k * P.a_sk;
            return FillSrc<T>{ok ? src : st.ptab, ok, i * AS + k};
        });
    }
    __syncthreads();

    const int nstrips = (rows + 15) >> 4;
    static_assert(NS >= 1 && NS <= 4, "strip count");
#define ZGL_TILES(n)                                                                                  \
    {                                                                                                \
        if constexpr (X3) {                                                                          \
            if (center) stream_tiles_x3<n, S::kCenter>(st, P, As, AS, m0, rows, g0, gn, wave, lane); \
            else stream_tiles_x3<n, false>(st, P, As, AS, m0, rows, g0, gn, wave, lane);             \
        } else {                                                                                     \
            if (center) stream_tiles<n, S::kCenter>(st, P, As, AS, m0, rows, g0, gn, c0, tiles, wave, lane); \
            else stream_tiles<n, false>(st, P, As, AS, m0, rows, g0, gn, c0, tiles, wave, lane);             \
        }                                                                                            \
    }
    switch (nstrips) {
        case 1: ZGL_TILES(1); break;
        case 2: if constexpr (NS >= 2) ZGL_TILES(2); break;
        case 3: if constexpr (NS >= 3) ZGL_TILES(3); break;
        default: if constexpr (NS >= 4) ZGL_TILES(4); break;
    }
#undef ZGL_TILES
}

template <class S, int NS, bool PLM, bool X3 = false>
__global__ __launch_bounds__(kThreads, 3) void stream_kernel(S st, const int4* __restrict__ work, int nb, int tpw) {
    stream_body<S, NS, PLM, X3>(st, work[blockIdx.x], blockIdx.y, nb, tpw);
}

// Persistent form over many orders (ragged batches): workgroup b walks items
// [bounds[b], bounds[b + 1]), item = (group, work index, instance group).
template <class S>
struct StreamGroup {
    S st;
    const int4* work;
    int nb;
    int nwork;  // items of the work list
};

template <class S, int NS, bool PLM>
__global__ __launch_bounds__(kThreads, 3) void stream_persistent(const StreamGroup<S>* __restrict__ groups,
                                                                 const int4* __restrict__ items,
                                                                 const int* __restrict__ bounds) {
    const int w1 = bounds[blockIdx.x + 1];
    for (int w = bounds[blockIdx.x]; w < w1; ++w) {
        const int4 it = items[w];
        const StreamGroup<S>& gr = groups[it.x];
        const S st = gr.st;
        __syncthreads();  // the previous item is done with the LDS
        stream_body<S, NS, PLM, false>(st, gr.work[it.y], it.z, gr.nb);
    }
}

// Two workgroups of 80 KB fit one CU's 160 KB of LDS.
constexpr size_t kMaxStreamLds = 80 * 1024;

// Dynamic LDS above 64 KB must be allowed per kernel.
template <class F>
void allow_lds(F* kernel, size_t lds) {
    if (lds > 64 * 1024)
        check_hip(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                  "LDS attribute");
}

int group_of(int requested, int nb) {
    const int g = requested > 0 ? std::min(requested, kMaxGroup) : 4;
    return nb < g ? nb : g;
}

// Instances per workgroup of a streaming stage: the requested group, capped
// so that the group's operand and output spans stay within the 2 GiB reach of
// the kernels' 32-bit buffer offsets; 0 when one instance exceeds it.
template <class S>
int stream_group(const S& st, int requested, int nb) {
    const long span = std::max(st.b_stride(), st.c_stride()) * (long)sizeof(typename S::T);
    if (span >= (long)kOutOfRange) return 0;
    return (int)std::min<long>(group_of(requested, nb), ((long)kOutOfRange - 1) / span);
}

// group: instances per workgroup in the low four bits, or with kGroupTiles
// set, tiles per wave (the workgroups then split the flattened (instance,
// column) index of the batch into runs of 4 such tiles; not with x3)
template <int NS, class S>
bool launch_stream(S st, int kmax, const int4* work, int nwork, int nb, int group, hipStream_t s,
                   bool x3 = false) {
    using T = typename S::T;
    if constexpr (!std::is_same_v<T, float>) x3 = false;
    const int kp = x3 ? (kmax + 15) & ~15 : (kmax + 3) & ~3;
    st.as = kp + 1;  // odd stride: the 16 rows of an A fragment fall in distinct banks
    const size_t lds = (size_t)(16 * NS + (S::kCenter ? 1 : 0)) * st.as * sizeof(T);
    if (lds > kMaxStreamLds) return false;
    if (nwork == 0 || nb <= 0) return true;
    int tpw = 0, ngroups = 0;
    if ((group & kGroupTiles) && !x3 && (group & 15) > 0) {
        // a run of 4 tpw tiles touches at most `span` instances (every
        // problem has N = 2 ncol columns); their operand and output spans
        // must stay within the 32-bit offsets, else instance groups are used
        constexpr int TNW = 16 * SM<T>::VW;
        const long n = 2L * st.ncol, run = 4L * (group & 15) * TNW;
        const long span = std::min<long>((run + n - 1) / n + 1, nb);
        const long ibytes = std::max(st.b_stride(), st.c_stride()) * (long)sizeof(T);
        if (span * ibytes < (long)kOutOfRange) {
            tpw = group & 15;
            st.G = (int)span;  // not read by the kernels in this mode
            ngroups = (int)((nb * n + run - 1) / run);
        }
    }
    if (tpw == 0) {
        st.G = stream_group(st, group & 15, nb);
        if (st.G <= 0) return false;
        ngroups = (nb + st.G - 1) / st.G;
    }
    dim3 grid(nwork, ngroups);
    auto go = [&](auto kernel) {
        allow_lds(kernel, lds);
        kernel<<<grid, kThreads, lds, s>>>(st, work, nb, tpw);
    };
    if constexpr (std::is_same_v<T, float>) {
        if (x3) {
            if (st.ptab) go(&stream_kernel<S, NS, false, true>);
            else go(&stream_kernel<S, NS, true, true>);
            check_hip(hipGetLastError(), "stream launch");
            return true;
        }
    }
    if (st.ptab) go(&stream_kernel<S, NS, false>);
    else go(&stream_kernel<S, NS, true>);
    check_hip(hipGetLastError(), "stream launch");
    return true;
}

// ------------------------------------------------------------------ radial
// Replaces integrate_radial (include/zernike_glq_transformer.hpp:819-865) and
// sum_n (include/zernike_glq_transformer.hpp:867-899).
//
// Forward: c[n,l,m] = kappa_n sum_k R_nl(r_k) flm[l,m][k], n = l + 2i.
// A workgroup owns one l, one block of up to 16*NS rows i and G instances; the
// R block sits in LDS and the flm block of each instance (rows m = 0..l,
// contiguous) streams through registers, 16 m per tile. The contraction axis
// k is contiguous in both operands, so in MFMA step t (t < KV) of an
// iteration lane group g takes k = k0 + KV*g + t: one 16-byte load of
// flm[l,m][k0 + KV*g ..] holds the (re, im) pairs of its KV k values, and the
// A operand reads the same k from LDS.
template <class T> struct RadCfg;
template <> struct RadCfg<float> { static constexpr int KV = 2; };   // float4 = 2 complex
template <> struct RadCfg<double> { static constexpr int KV = 1; };  // double2 = 1 complex

template <class T>
struct RadFwdArgs {
    const T* rtab; const T* flm; Inst<double> coeffs;
    long flm_stride;  // elements of T
    int nt, nr, unnormed, G, as;
};

template <int NSA, class T>
__device__ __forceinline__ void rad_fwd_tiles(const RadFwdArgs<T>& a, const T* As, double* const* cbase, int l,
                                              int i0, int rows, int g0, int gn, int wave, int lane) {
    using M = SM<T>;
    using V = typename M::V;
    constexpr int KV = RadCfg<T>::KV, KI = 4 * KV;
    const int j = lane & 15, g = lane >> 4;
    const int AS = a.as;
    const int iters = (a.nr + KI - 1) / KI;
    const int nmb = (l + 16) >> 4;  // blocks of 16 m in [0, l]
    const int tiles = gn * nmb;
    const int ntw = tiles > wave ? (tiles - wave + 3) >> 2 : 0;
    const int nchunks = (iters + kKC - 1) / kKC;
    const int U = ntw * nchunks;
    const unsigned rowbytes = 2u * a.nr * sizeof(T);
    const unsigned fbytes = (unsigned)(a.flm_stride * sizeof(T));
    const unsigned cbytes = (unsigned)tet_idx(a.nt, 0, 0) * 16u;  // coefficients n < nt

    typename M::acc_t are[NSA], aim[NSA];
#pragma unroll
    for (int q = 0; q < NSA; ++q) are[q] = aim[q] = typename M::acc_t{0, 0, 0, 0};

    auto unit = [&](int u, int& gi, int& m, int& ch) {
        const int tw = u / nchunks;
        ch = u - tw * nchunks;
        const int t = wave + 4 * tw;
        gi = t / nmb;
        m = 16 * (t - gi * nmb) + j;
    };
    // buffer loads: lanes past the block (m > l) or the k axis load zeros
    auto load = [&](V (&buf)[kKC], int u) {
        int gi, m, ch;
        unit(u, gi, m, ch);
        const __amdgpu_buffer_rsrc_t rf = buffer_range(a.flm + (g0 + gi) * a.flm_stride, fbytes);
        const unsigned rowoff = (unsigned)tri_idx(l, m) * rowbytes;
#pragma unroll
        for (int s = 0; s < kKC; ++s) {
            const int k = (ch * kKC + s) * KI + KV * g;
            const bool ok = m <= l && k < a.nr;
            buf[s] = buffer_load16<V>(rf, ok ? rowoff + 2u * k * sizeof(T) : kOutOfRange);
        }
    };
    // Every chunk issues the tile's coefficient stores -- dropped (out of
    // range) except after its last chunk -- so the count of stores between a
    // prefetch and its use is static and its wait never covers the stores.
    auto compute = [&](const V (&buf)[kKC], int u) {
        int gi, m, ch;
        unit(u, gi, m, ch);
#pragma unroll
        for (int s = 0; s < kKC; ++s) {
            const int it = ch * kKC + s;
            if (it < iters) {
#pragma unroll
                for (int t = 0; t < KV; ++t) {
#pragma unroll
                    for (int q = 0; q < NSA; ++q) {
                        const T x = As[(16 * q + j) * AS + it * KI + KV * g + t];
                        are[q] = M::mma(x, buf[s][2 * t], are[q]);
                        aim[q] = M::mma(x, buf[s][2 * t + 1], aim[q]);
                    }
                }
            }
        }
        const bool last = ch == nchunks - 1;
        const __amdgpu_buffer_rsrc_t rc = buffer_range(cbase[gi], cbytes);
#pragma unroll
        for (int q = 0; q < NSA; ++q) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * q + M::row(lane, r);
                const int n = l + 2 * (i0 + i);
                const double kappa = a.unnormed ? (double)(2 * n + 3) : 1.0;
                const bool ok = last && i < rows && m <= l;
                const d2 o = {(double)are[q][r] * kappa, (double)aim[q][r] * kappa};
                buffer_store16(rc, ok ? (unsigned)tet_idx(n, l, m) * 16u : kOutOfRange, o);
            }
            if (last) are[q] = aim[q] = typename M::acc_t{0, 0, 0, 0};
        }
    };

    if (U <= 0) return;
    V b0[kKC], b1[kKC];
    load(b0, 0);
    for (int u = 0; u < U; u += 2) {
        load(b1, min(u + 1, U - 1));
        compute(b0, u);
        if (u + 1 >= U) break;
        load(b0, min(u + 2, U - 1));
        compute(b1, u + 1);
    }
}

// FP64 operands need more registers than three waves per SIMD allow.
template <class T> constexpr int kRadWaves = sizeof(T) == 8 ? 2 : 3;

// bx: the workgroup index of the one-dimensional launch
template <class T, int NS>
__device__ __forceinline__ void rad_fwd_body(const RadFwdArgs<T>& a, const int4* __restrict__ work, const int bx,
                                             const int nb) {
    constexpr int TM = 16 * NS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* As = reinterpret_cast<T*>(smem_raw);
    // one-dimensional grid, instance group fastest: every group starts the
    // costliest items (the list is sorted by cost) first
    const int ng = (nb + a.G - 1) / a.G;
    const int wi = bx / ng;
    const int4 wk = work[wi];
    const int l = wk.x, i0 = wk.y;
    const int M = l < a.nt ? (a.nt - l + 1) / 2 : 0;
    const int rows = min(TM, M - i0);
    if (rows <= 0) return;
    const int g0 = (bx - wi * ng) * a.G;
    const int gn = min(a.G, nb - g0);
    if (gn <= 0) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int AS = a.as;  // >= KP: columns past nr are zeros
    lds_fill<8>(As, TM * AS, [&](int e) {
        const int i = e / AS, k = e - i * AS;
        const bool ok = i < rows && k < a.nr;
        return FillSrc<T>{a.rtab + (ok ? rad_idx(l + 2 * (i0 + i), l) * a.nr + k : 0L), ok, e};
    });
    // coefficient bases of the group's instances, looked up once
    __shared__ double* cbase[kMaxGroup];
    if (tid < gn) cbase[tid] = a.coeffs.at(g0 + tid);
    __syncthreads();
    switch ((rows + 15) >> 4) {
        case 1: rad_fwd_tiles<1>(a, As, cbase, l, i0, rows, g0, gn, wave, lane); break;
        case 2: if constexpr (NS >= 2) rad_fwd_tiles<2>(a, As, cbase, l, i0, rows, g0, gn, wave, lane); break;
        case 3: if constexpr (NS >= 3) rad_fwd_tiles<3>(a, As, cbase, l, i0, rows, g0, gn, wave, lane); break;
        default: if constexpr (NS >= 4) rad_fwd_tiles<4>(a, As, cbase, l, i0, rows, g0, gn, wave, lane); break;
    }
}

template <class T, int NS>
__global__ __launch_bounds__(kThreads, kRadWaves<T>) void rad_fwd_kernel(RadFwdArgs<T> a, const int4* __restrict__ work,
                                                             int nb) {
    rad_fwd_body<T, NS>(a, work, blockIdx.x, nb);
}

// Backward: flm[l,m][k] = sum_i R_nl(r_k) c[n,l,m], n = l + 2i.
// A workgroup owns one l and G instances; the R block (rows i, contiguous k)
// sits in LDS. A wave's tile is one instance and 16 values of m: it holds the
// (re, im) coefficients of its 16 m for every i in registers (the A operand)
// and sweeps the k axis in tiles of 16*VW columns with the column permutation
// of the Legendre kernels, storing VW interleaved (re, im) pairs per lane.
constexpr int kRadBwdSteps = 16;  // i < 64; the R block holds 16 * ceil(K / 16) rows

template <class T>
struct RadBwdArgs {
    const T* rtab; Inst<const double> coeffs; T* flm;
    long flm_stride;
    int nt, nr, G, rs;
    const int4* lorder = nullptr;  // l of rank r (nt entries, costliest first), or null
};

// The tiles of one wave with SC*4 MFMA steps (a compile-time count, so the
// coefficient loads and the MFMA sequence are branch-free).
template <int SC, class T>
__device__ __forceinline__ void rad_bwd_tiles(const RadBwdArgs<T>& a, const T* Rs, const double* const* cbase,
                                              int l, int K, int g0, int gn, int wave, int lane) {
    using M = SM<T>;
    using V = typename M::V;
    constexpr int VW = M::VW, TNW = 16 * VW;
    constexpr int NSTEP = 4 * SC;
    const int steps = (K + 3) >> 2;  // MFMA steps (the loads cover NSTEP)
    const int j = lane & 15, g = lane >> 4;
    const int RS = a.rs;
    const int nmb = (l + 16) >> 4;
    const int tiles = gn * nmb;
    if (wave >= tiles) return;
    const int ntile = (a.nr + TNW - 1) / TNW;  // <= NT: nr <= 136 (order <= 128)
    constexpr int NT = (136 + TNW - 1) / TNW;
    const unsigned rowbytes = 2u * a.nr * sizeof(T);
    const unsigned cbytes = (unsigned)tet_idx(a.nt, 0, 0) * 16u;  // coefficients n < nt
    const unsigned fbytes = (unsigned)(a.flm_stride * sizeof(T));

    // A operand of tile t: lane (row j -> m = mb + j, group g -> i = 4s + g),
    // the (re, im) FP64 pairs of its coefficients; lanes past the block load
    // zeros (out-of-range offsets)
    d2 cv[NSTEP];
    auto load = [&](int t) {
        const int gi = t / nmb;
        const int mb = 16 * (t - gi * nmb);
        const __amdgpu_buffer_rsrc_t rc = buffer_range(cbase[gi], cbytes);
#pragma unroll
        for (int s = 0; s < NSTEP; ++s) {
            const int i = 4 * s + g, m = mb + j;
            const bool ok = i < K && m <= l;
            cv[s] = buffer_load16<d2>(rc, ok ? (unsigned)tet_idx(l + 2 * i, l, m) * 16u : kOutOfRange);
        }
    };

    load(wave);
    for (int t = wave; t < tiles; t += 4) {
        T xre[NSTEP], xim[NSTEP];
#pragma unroll
        for (int s = 0; s < NSTEP; ++s) {
            xre[s] = (T)cv[s][0];
            xim[s] = (T)cv[s][1];
        }
        // the next tile's coefficients are requested before this tile's
        // stores, so waiting for them never waits for the stores (past the
        // last tile this re-reads the current one)
        load(t + 4 < tiles ? t + 4 : t);
        const int gi = t / nmb;
        const int mb = 16 * (t - gi * nmb);
        const __amdgpu_buffer_rsrc_t rf = buffer_range(a.flm + (g0 + gi) * a.flm_stride, fbytes);
        // a static count of column tiles (the stores of tiles past ntile are
        // dropped), so the count of stores after the prefetch is static too
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
            // A last tile of at most 16 columns (8 of 64 at order 128 in
            // FP32) takes one column per lane: 2 MFMAs per step instead of
            // 2 VW, and as many stores as a full tile (4 of them dropped)
            if (ct < ntile && a.nr - ct * TNW <= 16) {
                typename M::acc_t ar1 = {0, 0, 0, 0}, ai1 = {0, 0, 0, 0};
                const int col = ct * TNW + j;
#pragma unroll
                for (int s = 0; s < NSTEP; ++s) {
                    if (s < steps) {
                        const T b = Rs[(4 * s + g) * RS + col];
                        ar1 = M::mma(xre[s], b, ar1);
                        ai1 = M::mma(xim[s], b, ai1);
                    }
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = mb + M::row(lane, r);
                    const bool ok = m <= l && col < a.nr;
                    const unsigned o = (unsigned)tri_idx(l, m) * rowbytes + 2u * col * sizeof(T);
                    buffer_store_pair<T>(rf, ok ? o : kOutOfRange, ar1[r], ai1[r]);
                    buffer_store16(rf, kOutOfRange, V{});
                }
                continue;
            }
            typename M::acc_t are[VW], aim[VW];
#pragma unroll
            for (int v = 0; v < VW; ++v) are[v] = aim[v] = typename M::acc_t{0, 0, 0, 0};
            // column of value v of lane j: half q = v / HV of the tile, then
            // HV consecutive columns per lane, so that the 16-byte (re, im)
            // store of each half is contiguous across the 16 lanes of a row
            // (256 bytes per store instruction and row, not 16-byte pieces
            // 32 bytes apart)
            constexpr int HV = VW / 2;
            const int col0 = ct * TNW + HV * j;
#pragma unroll
            for (int s = 0; s < NSTEP; ++s) {
                if (s < steps && ct < ntile) {
                    const T* rrow = Rs + (4 * s + g) * RS + col0;
                    T b[VW];
#pragma unroll
                    for (int q = 0; q < 2; ++q)
#pragma unroll
                        for (int e = 0; e < HV; ++e) b[q * HV + e] = rrow[q * (TNW / 2) + e];
#pragma unroll
                    for (int v = 0; v < VW; ++v) {
                        are[v] = M::mma(xre[s], b[v], are[v]);
                        aim[v] = M::mma(xim[s], b[v], aim[v]);
                    }
                }
            }
            // per row and half: the (re, im) pairs of HV columns, one 16-byte
            // store; rows past l and columns past nr are dropped
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = mb + M::row(lane, r);
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int c = col0 + h * (TNW / 2);
                    const bool ok = m <= l && c < a.nr;
                    const unsigned o = (unsigned)tri_idx(l, m) * rowbytes + 2u * c * sizeof(T);
                    V w;
#pragma unroll
                    for (int e = 0; e < HV; ++e) {
                        w[2 * e] = are[h * HV + e][r];
                        w[2 * e + 1] = aim[h * HV + e][r];
                    }
                    buffer_store16(rf, ok ? o : kOutOfRange, w);
                }
            }
        }
    }
}

template <class T>
__device__ __forceinline__ void rad_bwd_body(const RadBwdArgs<T>& a, const int bx, const int nb) {
    constexpr int TNW = 16 * SM<T>::VW;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* Rs = reinterpret_cast<T*>(smem_raw);
    // one-dimensional grid, instance group fastest; the rank r enumerates l
    // from the middle outwards (c, c-1, c+1, ..., c = nt/2), roughly by
    // decreasing cost (l+1)(nt-l+1)/2, and covers [0, nt) exactly once
    // (or, given, from the list of l by decreasing cost)
    const int ng = (nb + a.G - 1) / a.G;
    const int r = bx / ng, d = (r + 1) >> 1;
    const int l = a.lorder ? a.lorder[r].x : (r & 1) ? a.nt / 2 - d : a.nt / 2 + d;
    if (l < 0 || l >= a.nt) return;
    const int K = (a.nt - l + 1) / 2;
    const int sc = (K + 15) >> 4;  // groups of four MFMA steps, K <= 4 * kRadBwdSteps
    const int g0 = (bx - r * ng) * a.G;
    const int gn = min(a.G, nb - g0);
    if (gn <= 0) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int RS = a.rs;
    // Rs[i * RS + k], RS = nr, rows i < 16 sc (zero past K); columns of the
    // last (partial) tile past nr read the next row or the zero tail, and
    // their results are not stored
    lds_fill<8>(Rs, 16 * sc * RS + TNW, [&](int e) {
        const int i = e / RS, k = e - i * RS;
        const bool ok = i < K && k < a.nr;
        return FillSrc<T>{a.rtab + (ok ? rad_idx(l + 2 * i, l) * a.nr + k : 0L), ok, e};
    });
    // coefficient bases of the group's instances, looked up once (a ragged
    // batch reads them from its offset table), so the tile loop has no
    // dependent global load
    __shared__ const double* cbase[kMaxGroup];
    if (tid < gn) cbase[tid] = a.coeffs.at(g0 + tid);
    __syncthreads();
    switch (sc) {
        case 1: rad_bwd_tiles<1>(a, Rs, cbase, l, K, g0, gn, wave, lane); break;
        case 2: rad_bwd_tiles<2>(a, Rs, cbase, l, K, g0, gn, wave, lane); break;
        case 3: rad_bwd_tiles<3>(a, Rs, cbase, l, K, g0, gn, wave, lane); break;
        default: rad_bwd_tiles<4>(a, Rs, cbase, l, K, g0, gn, wave, lane); break;
    }
}

// The backward keeps the next tile's coefficients in flight beside the
// current tile's: two waves per SIMD.
constexpr int kRadBwdWaves = 2;

template <class T>
__global__ __launch_bounds__(kThreads, kRadBwdWaves) void rad_bwd_kernel(RadBwdArgs<T> a, int nb) {
    rad_bwd_body<T>(a, blockIdx.x, nb);
}

// Persistent forms of the radial kernels (ragged batches): item = (group,
// workgroup index of the group's one-dimensional launch).
template <class T>
struct RadFwdGroup {
    RadFwdArgs<T> a;
    const int4* work;
    int nb;
};
template <class T>
struct RadBwdGroup {
    RadBwdArgs<T> a;
    int nb;
};

template <class T, int NS>
__global__ __launch_bounds__(kThreads, kRadWaves<T>) void rad_fwd_persistent(const RadFwdGroup<T>* __restrict__ groups,
                                                                  const int4* __restrict__ items,
                                                                  const int* __restrict__ bounds) {
    const int w1 = bounds[blockIdx.x + 1];
    for (int w = bounds[blockIdx.x]; w < w1; ++w) {
        const int4 it = items[w];
        const RadFwdGroup<T>& gr = groups[it.x];
        const RadFwdArgs<T> a = gr.a;
        __syncthreads();
        rad_fwd_body<T, NS>(a, gr.work, it.y, gr.nb);
    }
}

template <class T>
__global__ __launch_bounds__(kThreads, kRadBwdWaves) void rad_bwd_persistent(const RadBwdGroup<T>* __restrict__ groups,
                                                                  const int4* __restrict__ items,
                                                                  const int* __restrict__ bounds) {
    const int w1 = bounds[blockIdx.x + 1];
    for (int w = bounds[blockIdx.x]; w < w1; ++w) {
        const int4 it = items[w];
        const RadBwdGroup<T>& gr = groups[it.x];
        const RadBwdArgs<T> a = gr.a;
        __syncthreads();
        rad_bwd_body<T>(a, it.y, gr.nb);
    }
}

template <class T>
const T* ptab_sel(const OrderTables& t);
template <> const double* ptab_sel<double>(const OrderTables& t) { return t.ptab; }
template <> const float* ptab_sel<float>(const OrderTables& t) { return t.ptab32; }

size_t stream_lds(int rows_tm, bool center, int kmax, size_t tsize) {
    const int kp = (kmax + 3) & ~3;
    return (size_t)(rows_tm + (center ? 1 : 0)) * (kp + 1) * tsize;
}

}  // namespace

template <class T>
const T* rtab_sel(const OrderTables& t);
template <> const double* rtab_sel<double>(const OrderTables& t) { return t.rtab; }
template <> const float* rtab_sel<float>(const OrderTables& t) { return t.rtab32; }

template <class T>
bool zt_radial_forward_stream(const ZtGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    constexpr int KI = 4 * RadCfg<T>::KV;
    const int kp = (g.ncol + KI - 1) / KI * KI;
    RadFwdArgs<T> r{rtab_sel<T>(*a.t), (const T*)a.flm, a.coeffs, a.flm_stride * 2,
                    a.nt, g.ncol, a.zernike_unnormed, 1, kp + 2};
    const size_t lds = (size_t)16 * kNsRadial * r.as * sizeof(T);
    if (lds > kMaxStreamLds) return false;
    if ((size_t)a.flm_stride * 2 * sizeof(T) >= kOutOfRange) return false;  // buffer ranges
    if (a.t->n_rf == 0 || a.nb <= 0) return true;
    r.G = group_of(a.group, a.nb);
    dim3 grid((unsigned)a.t->n_rf * (unsigned)((a.nb + r.G - 1) / r.G));
    allow_lds(&rad_fwd_kernel<T, kNsRadial>, lds);
    rad_fwd_kernel<T, kNsRadial><<<grid, kThreads, lds, s>>>(r, a.t->work_rf, a.nb);
    check_hip(hipGetLastError(), "radial forward launch");
    return true;
}

template <class T>
bool zt_radial_backward_stream(const ZtGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    constexpr int TNW = 16 * SM<T>::VW;
    const int kmax = (g.order + 1) / 2;  // i values of l = 0
    if (kmax > 4 * kRadBwdSteps || g.ncol > 136) return false;
    // rows of ncol reals, plus one column tile of zeros so that the last
    // tile of the last row reads inside the array
    const int rs = g.ncol;
    const size_t lds = ((size_t)16 * ((kmax + 15) / 16) * rs + TNW) * sizeof(T);
    if (lds > kMaxStreamLds) return false;
    if (a.nt == 0 || a.nb <= 0) return true;
    RadBwdArgs<T> r{rtab_sel<T>(*a.t), a.coeffs, (T*)a.flm, a.flm_stride * 2,
                    a.nt, g.ncol, group_of(a.group, a.nb), rs};
    // the cost-ordered list covers l < order: used when nothing is truncated
    if (a.t->n_rs == a.nt) r.lorder = a.t->work_rs;
    dim3 grid((unsigned)a.nt * (unsigned)((a.nb + r.G - 1) / r.G));
    allow_lds(&rad_bwd_kernel<T>, lds);
    rad_bwd_kernel<T><<<grid, kThreads, lds, s>>>(r, a.nb);
    check_hip(hipGetLastError(), "radial backward launch");
    return true;
}

// ZtGemmArgs::group of the Legendre stages: instances per workgroup in the
// low four bits, kGroupX3 set for the BF16x3 products (FP32 stages);
// ZGL_LEGENDRE_MATH=bf16x3 selects them when the caller leaves it unset.
static bool legendre_x3(int group) {
    if (group & kGroupX3) return true;
    if (group != 0) return false;
    const char* e = std::getenv("ZGL_LEGENDRE_MATH");
    return e && std::strcmp(e, "bf16x3") == 0;
}

// The grouping of an untuned Legendre stage: ZGL_LEGENDRE_GROUP forces one
// (1..16 instances per workgroup, or 33..47: kGroupTiles | tiles per wave),
// else the default (4 instances).
static int legendre_group(int group) {
    if (group & ~kGroupX3) return group & ~kGroupX3;
    const char* e = std::getenv("ZGL_LEGENDRE_GROUP");
    if (!e || !*e) return 0;
    const int v = std::atoi(e);
    return (v >= 1 && v <= kMaxGroup) || (v > kGroupTiles && v < kGroupTiles + 16) ? v : 0;
}

template <class T>
bool zt_legendre_forward_stream(const ZtGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    LegFwd<T> st{ptab_sel<T>(*a.t), (const T*)a.spec, (T*)a.flm, a.spec_stride * 2,
                 a.flm_stride * 2, g.order, a.nt, g.jh, g.ncol, 1, 0,
                 a.t->zh, a.t->plm_qm, a.t->plm_phase};
    return launch_stream<kNsLegendreFwd>(st, g.jh, a.t->work_lf, a.t->n_lf, a.nb, legendre_group(a.group), s,
                                         legendre_x3(a.group));
}

template <class T>
bool zt_legendre_backward_stream(const ZtGemmArgs& a, hipStream_t s) {
    const Geom& g = a.t->g;
    LegBwd<T> st{ptab_sel<T>(*a.t), (const T*)a.flm, (T*)a.spec, a.spec_stride * 2,
                 a.flm_stride * 2, g.order, a.nt, g.jh, g.ncol, 1, 0, g.nth,
                 a.t->zh, a.t->plm_qm, a.t->plm_phase};
    return launch_stream<kNsStreamBwd>(st, plm_count(g.order, 0, 0), a.t->work_sb, a.t->n_sb,
                                       a.nb, legendre_group(a.group), s, legendre_x3(a.group));
}

// ---- persistent launches over many orders (ragged batches) ----
namespace {

// Uploads the group array and the dealt items of one persistent launch and
// appends it to out.
template <class Group, class Kernel>
void persistent_launch(std::vector<PersistentLaunch>& out, const char* stage, Kernel* k, size_t lds,
                       const std::vector<Group>& groups, const std::vector<int4>& items,
                       const std::vector<double>& cost) {
    if (items.empty()) return;
    PersistentLaunch pl;
    pl.stage = stage;
    for (const Group& g : groups) pl.instances += g.nb;
    allow_lds(k, lds);
    std::vector<int4> dealt;
    std::vector<int> bounds;
    lpt_deal(items, cost, resident_blocks(reinterpret_cast<const void*>(k), kThreads, lds), dealt, bounds, pl);
    const int blocks = (int)bounds.size() - 1;
    const Group* dg = upload_owned(pl, groups);
    const int4* di = upload_owned(pl, dealt);
    const int* db = upload_owned(pl, bounds);
    pl.items = (long)items.size();
    pl.blocks = blocks;
    pl.run = [=](hipStream_t s) {
        k<<<blocks, kThreads, lds, s>>>(dg, di, db);
        check_hip(hipGetLastError(), stage);
    };
    out.push_back(std::move(pl));
}

// Items (group, work index, instance group) of a streaming Legendre stage;
// the modelled cost of an item is its full tile: 16 NS rows x kmax x ncol x G.
template <int NS, class S>
bool legendre_persistent(std::vector<PersistentLaunch>& out, const char* stage, std::vector<StreamGroup<S>>& groups,
                         const std::vector<int>& kmax) {
    using T = typename S::T;
    size_t lds = 0;
    std::vector<int4> items;
    std::vector<double> cost;
    for (int j = 0; j < (int)groups.size(); ++j) {
        StreamGroup<S>& gr = groups[j];
        if (!gr.st.ptab) return false;  // P blocks recomputed in the kernel: no persistent form
        const int kp = (kmax[j] + 3) & ~3;
        gr.st.as = kp + 1;
        const size_t l = (size_t)(16 * NS + (S::kCenter ? 1 : 0)) * gr.st.as * sizeof(T);
        if (l > kMaxStreamLds) return false;
        lds = std::max(lds, l);
    }
    for (int j = 0; j < (int)groups.size(); ++j) {
        const StreamGroup<S>& gr = groups[j];
        const int ng = (gr.nb + gr.st.G - 1) / gr.st.G;
        const double c = 16.0 * NS * kmax[j] * gr.st.ncol * gr.st.G;
        for (int w = 0; w < gr.nwork; ++w)
            for (int ig = 0; ig < ng; ++ig) {
                items.push_back(make_int4(j, w, ig, 0));
                cost.push_back(c);
            }
    }
    persistent_launch(out, stage, &stream_persistent<S, NS, false>, lds, groups, items, cost);
    return true;
}

}  // namespace

template <class T>
bool zt_legendre_forward_persistent(std::vector<PersistentLaunch>& out, const std::vector<ZtGemmArgs>& g) {
    std::vector<StreamGroup<LegFwd<T>>> groups;
    std::vector<int> kmax;
    for (const ZtGemmArgs& a : g) {
        const Geom& G = a.t->g;
        LegFwd<T> st{ptab_sel<T>(*a.t), (const T*)a.spec, (T*)a.flm, a.spec_stride * 2,
                     a.flm_stride * 2, G.order, a.nt, G.jh, G.ncol, 0, 0,
                     a.t->zh, a.t->plm_qm, a.t->plm_phase};
        st.G = stream_group(st, a.group & 15, a.nb);
        if (st.G <= 0) return false;
        groups.push_back({st, a.t->work_lf, a.nb, a.t->n_lf});
        kmax.push_back(G.jh);
    }
    return legendre_persistent<kNsLegendreFwd>(out, "legendre_forward", groups, kmax);
}

template <class T>
bool zt_legendre_backward_persistent(std::vector<PersistentLaunch>& out, const std::vector<ZtGemmArgs>& g) {
    std::vector<StreamGroup<LegBwd<T>>> groups;
    std::vector<int> kmax;
    for (const ZtGemmArgs& a : g) {
        const Geom& G = a.t->g;
        LegBwd<T> st{ptab_sel<T>(*a.t), (const T*)a.flm, (T*)a.spec, a.spec_stride * 2,
                     a.flm_stride * 2, G.order, a.nt, G.jh, G.ncol, 0, 0, G.nth,
                     a.t->zh, a.t->plm_qm, a.t->plm_phase};
        st.G = stream_group(st, a.group & 15, a.nb);
        if (st.G <= 0) return false;
        groups.push_back({st, a.t->work_sb, a.nb, a.t->n_sb});
        kmax.push_back(plm_count(G.order, 0, 0));
    }
    return legendre_persistent<kNsStreamBwd>(out, "legendre_backward", groups, kmax);
}

template <class T>
bool zt_radial_forward_persistent(std::vector<PersistentLaunch>& out, const std::vector<ZtGemmArgs>& g) {
    constexpr int KI = 4 * RadCfg<T>::KV;
    std::vector<RadFwdGroup<T>> groups;
    std::vector<int4> items;
    std::vector<double> cost;
    size_t lds = 0;
    for (int j = 0; j < (int)g.size(); ++j) {
        const ZtGemmArgs& a = g[j];
        const Geom& G = a.t->g;
        const int kp = (G.ncol + KI - 1) / KI * KI;
        RadFwdArgs<T> r{rtab_sel<T>(*a.t), (const T*)a.flm, a.coeffs, a.flm_stride * 2,
                        a.nt, G.ncol, a.zernike_unnormed, group_of(a.group, a.nb), kp + 2};
        const size_t l = (size_t)16 * kNsRadial * r.as * sizeof(T);
        if (l > kMaxStreamLds) return false;
        lds = std::max(lds, l);
        groups.push_back({r, a.t->work_rf, a.nb});
        const int ng = (a.nb + r.G - 1) / r.G;
        // an item: up to 16 NS rows of one l against its l + 1 values of m, G instances
        const double c = 16.0 * kNsRadial * G.ncol * (0.5 * a.nt + 1) * r.G;
        for (int bx = 0; bx < a.t->n_rf * ng; ++bx) {
            items.push_back(make_int4(j, bx, 0, 0));
            cost.push_back(c);
        }
    }
    persistent_launch(out, "radial_forward", &rad_fwd_persistent<T, kNsRadial>, lds, groups, items, cost);
    return true;
}

template <class T>
bool zt_radial_backward_persistent(std::vector<PersistentLaunch>& out, const std::vector<ZtGemmArgs>& g) {
    constexpr int TNW = 16 * SM<T>::VW;
    std::vector<RadBwdGroup<T>> groups;
    std::vector<int4> items;
    std::vector<double> cost;
    size_t lds = 0;
    for (int j = 0; j < (int)g.size(); ++j) {
        const ZtGemmArgs& a = g[j];
        const Geom& G = a.t->g;
        const int kmax = (G.order + 1) / 2;
        if (kmax > 4 * kRadBwdSteps || G.ncol > 136) return false;
        const int rs = G.ncol;
        const size_t l = ((size_t)16 * ((kmax + 15) / 16) * rs + TNW) * sizeof(T);
        if (l > kMaxStreamLds) return false;
        lds = std::max(lds, l);
        RadBwdArgs<T> r{rtab_sel<T>(*a.t), a.coeffs, (T*)a.flm, a.flm_stride * 2,
                        a.nt, G.ncol, group_of(a.group, a.nb), rs};
        groups.push_back({r, a.nb});
        const int ng = (a.nb + r.G - 1) / r.G;
        for (int bx = 0; bx < a.nt * ng; ++bx) {
            // rank r of the kernel's middle-out order of l, as rad_bwd_body maps it
            const int rk = bx / ng, d = (rk + 1) >> 1;
            const int lv = (rk & 1) ? a.nt / 2 - d : a.nt / 2 + d;
            if (lv < 0 || lv >= a.nt) continue;
            items.push_back(make_int4(j, bx, 0, 0));
            cost.push_back((double)(lv + 1) * ((a.nt - lv + 1) / 2) * G.ncol * r.G);
        }
    }
    persistent_launch(out, "radial_backward", &rad_bwd_persistent<T>, lds, groups, items, cost);
    return true;
}

bool stream_legendre_fits(int order, int jh) {
    return stream_lds(16 * kNsLegendreFwd, false, jh, sizeof(double)) <= kMaxStreamLds &&
           stream_lds(16 * kNsStreamBwd, true, plm_count(order, 0, 0), sizeof(double)) <= kMaxStreamLds;
}

// Longest l range of one pass: the P block of one 16-point strip fills the
// LDS budget (80 KB / (16 * 8 B) = 640 values per row, less the padding).
constexpr int kEvalKChunk = 636;

int eval_legendre_strips(int nt) {
    for (int ns : {4, 2})
        if (stream_lds(16 * ns, false, nt, sizeof(double)) <= kMaxStreamLds) return ns;
    return 1;
}

bool eval_legendre(const EvalLegArgs& a, hipStream_t s) {
    const int kchunk = a.ns == 1 ? std::min(a.nt, kEvalKChunk) : a.nt;
    for (int l0 = 0; l0 < a.nt; l0 += kchunk) {
        LegEval<double> st{a.ptab, a.flm, a.spec, a.spec_stride, a.flm_stride, a.order, a.nt, a.nlat,
                           a.ncol, 1, 0, a.z, a.qm, a.phase, l0, kchunk, l0 > 0 ? 1 : 0};
        bool ok = false;
        switch (a.ns) {
            case 4: ok = launch_stream<4>(st, kchunk, a.work, a.nwork, a.nb, 0, s); break;
            case 2: ok = launch_stream<2>(st, kchunk, a.work, a.nwork, a.nb, 0, s); break;
            case 1: ok = launch_stream<1>(st, kchunk, a.work, a.nwork, a.nb, 0, s); break;
            default: break;
        }
        if (!ok) return false;
    }
    return true;
}

template bool zt_legendre_forward_stream<double>(const ZtGemmArgs&, hipStream_t);
template bool zt_legendre_forward_stream<float>(const ZtGemmArgs&, hipStream_t);
template bool zt_legendre_backward_stream<double>(const ZtGemmArgs&, hipStream_t);
template bool zt_legendre_backward_stream<float>(const ZtGemmArgs&, hipStream_t);
template bool zt_radial_forward_stream<double>(const ZtGemmArgs&, hipStream_t);
template bool zt_radial_forward_stream<float>(const ZtGemmArgs&, hipStream_t);
template bool zt_radial_backward_stream<double>(const ZtGemmArgs&, hipStream_t);
template bool zt_radial_backward_stream<float>(const ZtGemmArgs&, hipStream_t);
template bool zt_legendre_forward_persistent<double>(std::vector<PersistentLaunch>&, const std::vector<ZtGemmArgs>&);
template bool zt_legendre_forward_persistent<float>(std::vector<PersistentLaunch>&, const std::vector<ZtGemmArgs>&);
template bool zt_legendre_backward_persistent<double>(std::vector<PersistentLaunch>&, const std::vector<ZtGemmArgs>&);
template bool zt_legendre_backward_persistent<float>(std::vector<PersistentLaunch>&, const std::vector<ZtGemmArgs>&);
template bool zt_radial_forward_persistent<double>(std::vector<PersistentLaunch>&, const std::vector<ZtGemmArgs>&);
template bool zt_radial_forward_persistent<float>(std::vector<PersistentLaunch>&, const std::vector<ZtGemmArgs>&);
template bool zt_radial_backward_persistent<double>(std::vector<PersistentLaunch>&, const std::vector<ZtGemmArgs>&);
template bool zt_radial_backward_persistent<float>(std::vector<PersistentLaunch>&, const std::vector<ZtGemmArgs>&);

}  // namespace zgl
