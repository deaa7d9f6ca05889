// Fused colatitude + radial contractions of the Zernike transforms, FP32.
//
// Replaces, for the FP32 modes at orders up to 128, the stage pairs
//   integrate_latitudinal + integrate_radial  include/zernike_glq_transformer.hpp:778-865
//   sum_n + sum_l                             include/zernike_glq_transformer.hpp:867-941
// without the flm intermediate in HBM:
//   forward   c[n,l,m]     = kappa_n sum_k R_nl(r_k) sum_h P_lm(z_h) X^par_m[h][k]
//   backward  X^par_m[h][k] = sum_l P_lm(z_h) sum_n R_nl(r_k) c[n,l,m]
// with l = m + par + 2i (the parity-split rows of the spectrum X) and n = l + 2d.
//
// A workgroup (8 waves) owns one (m, parity) problem and 8 instances, and
// walks the radial nodes in chunks of 16. The radial contraction is a GEMM
// per l whose columns are the (instance, re/im) pairs, so the 8 instances
// give it a full 16-column MFMA tile; the flm values of one chunk (all rows
// l of the problem, 16 nodes, 16 columns) pass from one contraction to the
// other through LDS, in two buffers so that consecutive chunks overlap and
// one barrier per chunk suffices.
//
// Forward: waves 0-3 produce, waves 4-7 consume. Producer p streams the
// spectrum rows of instances 2p and 2p+1 (16 nodes x re/im = 128 contiguous
// bytes per instance and half latitude, the next chunk's loads in flight
// during the current chunk's products) against the P block in LDS (rows l,
// at most 64) and writes its 64 x 64 flm tile into one buffer while the
// consumers contract the other buffer's chunk with R_nl(r_k) into radial
// accumulators that stay in registers over all chunks (rows i = c + 4r of
// the problem, tiles of 16 values of n: at most 40 per consumer wave). Both
// kinds of products share each SIMD's matrix core, one wave of each kind.
// Backward: in the iteration of chunk k every wave forms the flm values of
// its rows i = w + 8r for chunk k (MFMA rows = the 16 nodes, K = n, B = the
// coefficients, loaded once per workgroup) into one buffer, then contracts
// the P block (rows = half latitudes) against instance w's flm columns of
// chunk k - 1 from the other buffer and stores those spectrum rows.
//
// MFMA v_mfma_f32_16x16x4f32: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and receives D[4 (l / 16) + r][l % 16], r < 4. The K index
// of a step may be any permutation of the contraction axis as long as A and B
// use the same one: the radial steps take k = 4g + s (forward) or
// d = 16b + 4g + s (backward) in step s of lane group g, so one 16-byte load
// holds a lane's operands of four steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "zgl_internal.hpp"

namespace zgl {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));

constexpr int kFT = 512;          // threads per workgroup
constexpr int kFW = kFT / 64;     // waves = instances per workgroup
constexpr int kFK = 16;           // radial nodes per chunk
constexpr int kFRows = 64;        // rows l of one (m, parity) problem at most
constexpr int kFAS = 69;          // LDS row stride of the P block (odd: distinct banks)
constexpr int kFSteps = 17;       // 4-point steps over at most 68 half latitudes
// Forward flm buffer: Ft[i][k / 4][column][k % 4], strides kFQS and kFFS
// (padded so that the producers' scattered stores spread over the banks and
// the consumers' 16-byte loads stay aligned)
constexpr int kFQS = 68;
constexpr int kFFS = 276;
constexpr int kFtFwd = kFRows * kFFS;
// Backward flm buffer: Ft[i][instance][k ^ (2 instance)][re/im]
constexpr int kFtBwd = kFRows * kFW * kFK * 2;

// Forward radial tiles (16 values of n) of consumer row r: rows i = c + 4r
// have at most 64 - 4r values of n.
__host__ __device__ constexpr int ftiles4(int r) { return (kFRows - 4 * r + 15) / 16; }
__host__ __device__ constexpr int fslot4(int r) { return r == 0 ? 0 : fslot4(r - 1) + ftiles4(r - 1); }
constexpr int kFSlots4 = fslot4(16);  // 40
// Backward: 16-value blocks of n of wave-row r (rows i = w + 8r, four MFMA
// steps each)
__host__ __device__ constexpr int fblocks(int r) { return (kFRows - kFW * r + 15) / 16; }
__host__ __device__ constexpr int fslot(int r) { return r == 0 ? 0 : fslot(r - 1) + fblocks(r - 1); }
constexpr int kFSlots = fslot(kFW);  // 20

__device__ inline f4 mma(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// f(std::integral_constant<int, I>) for I = B .. E-1: row loops whose
// accumulator indices must be compile-time constants (register arrays)
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

struct FusedArgs {
    const float* ptab;   // P_lm(z_h), parity-split [m][par][i][h] (rows of the table's order)
    const float* rtab;   // forward: [l][kc][d][16 k]; backward: [l][kc][16 k][d], d < dp
    float* spec;         // spectrum of the chunk's instances (floats)
    long spec_stride;    // floats per instance
    Inst<double> coeffs;
    const int2* work;    // (m, par), costliest first
    int order, nt, jh, ncol, nth, nkc, dp, unnormed, nb, ngroups;
};

// ------------------------------------------------------------------ forward
// producer wave p: colatitude products of instances g0 + 2p, g0 + 2p + 1
template <int NSA>
__device__ __forceinline__ void fused_fwd_produce(const FusedArgs& a, const float* As, float* Ft0, float* Ft1,
                                                  int m, int par, int g0, int p, int lane) {
    const int j = lane & 15, g = lane >> 4;
    const int K = a.jh;
    const long rowlen = 2L * a.ncol;
    // Loads are unmasked, with clamped indices: rows h >= K meet zero P values,
    // nodes past ncol meet zero R values in the radial products, and an
    // instance past the batch is never stored, so every clamped value only
    // enters products that are discarded or zero. (No branch or mask separates
    // a load from its wait, so the waits count exactly.)
    const int inst = min(g0 + 2 * p + (j >> 3), a.nb - 1);
    const float* sb = a.spec + (long)inst * a.spec_stride + (long)(2 * m + par) * a.jh * rowlen;
    // lane columns: nodes 2 (j & 7) and 2 (j & 7) + 1 of the chunk, (re, im) each
    const int kl = 2 * (j & 7);
    auto ld = [&](int kc, int s) {
        const int k = min(kc * kFK + kl, a.ncol - 2);
        const int h = min(4 * s + g, K - 1);
        return *reinterpret_cast<const f4*>(sb + 2 * k + h * rowlen);
    };
    // this lane's P values (A operand) of step s, read from LDS one step ahead
    auto lda = [&](float (&x)[NSA], int s) {
#pragma unroll
        for (int q = 0; q < NSA; ++q) x[q] = As[(16 * q + j) * kFAS + 4 * s + g];
    };
    // one rolling register buffer: the register of step s is reloaded with the
    // next chunk's value as soon as the products of step s have read it, so
    // every load has a whole chunk of products to arrive
    f4 buf[kFSteps];
#pragma unroll
    for (int s = 0; s < kFSteps; ++s) buf[s] = ld(0, s);
    const int steps = (K + 3) >> 2;
    for (int kc = 0; kc < a.nkc; ++kc) {
        const int kn = min(kc + 1, a.nkc - 1);  // the last chunk reloads itself, unused
        f4 acc[NSA][4];
#pragma unroll
        for (int q = 0; q < NSA; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[q][v] = f4{0, 0, 0, 0};
        float ar[2][NSA];
        lda(ar[0], 0);
#pragma unroll
        for (int s = 0; s < kFSteps; ++s) {
            if (s + 1 < kFSteps) lda(ar[(s + 1) & 1], s + 1);
            if (s < steps) {
#pragma unroll
                for (int q = 0; q < NSA; ++q)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[q][v] = mma(ar[s & 1][q], buf[s][v], acc[q][v]);
            }
            buf[s] = ld(kn, s);
        }
        // column v of lane j: node kl + v/2, instance column 2 (2p + j/8) + v%2
        float* Ft = (kc & 1) ? Ft1 : Ft0;
        const int cbase = 2 * (2 * p + (j >> 3));
#pragma unroll
        for (int q = 0; q < NSA; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = 16 * q + 4 * g + r, k = kl + (v >> 1);
                    Ft[i * kFFS + (k >> 2) * kFQS + (cbase + (v & 1)) * 4 + (k & 3)] = acc[q][v][r];
                }
        __syncthreads();
    }
    __syncthreads();  // the consumers' last chunk
}

// consumer wave c: radial products of rows i = c + 4r, then the coefficients
__device__ __forceinline__ void fused_fwd_consume(const FusedArgs& a, const float* Ft0, const float* Ft1, int m,
                                                  int par, int M, int g0, int c, int lane) {
    const int j = lane & 15, g = lane >> 4;
    const int lbase = m + par;
    f4 racc[kFSlots4];
#pragma unroll
    for (int q = 0; q < kFSlots4; ++q) racc[q] = f4{0, 0, 0, 0};

    // Operands of row r: A = R rows n = l + 2 (16 t + j) at nodes 4 g .. 4 g + 3
    // of the chunk (rows past the table clamped to its last d: their products
    // are never formed), B = the chunk's flm values of row i. Rows past M load
    // row M - 1 and form no products. Loads are unmasked (see the producer).
    // (the table holds dp = 64 values of d for every l, so 16 t + j < 64 is in
    // it). 32-bit element offsets: a uniform row part plus the lane's part.
    const unsigned lstride = (unsigned)a.nkc * kFRows * kFK;  // floats per l in the table
    const unsigned lane_off = j * kFK + 4 * g;
    struct Row { f4 av[4]; f4 bv; };
    auto loadrow = [&](Row& w, auto rc, int kc, const float* Ft) {
        constexpr int r = decltype(rc)::value;
        const int i = min(c + 4 * r, M - 1);
        const unsigned row_off = (unsigned)(lbase + 2 * i) * lstride + (unsigned)kc * (kFRows * kFK);
        const float* rrow = a.rtab + (row_off + lane_off);
#pragma unroll
        for (int t = 0; t < ftiles4(r); ++t)
            w.av[t] = *reinterpret_cast<const f4*>(rrow + 16 * t * kFK);
        w.bv = *reinterpret_cast<const f4*>(Ft + i * kFFS + g * kFQS + j * 4);
    };
    // the products of row r, for its number of 16-row tiles of n (1..4)
    auto rowmma = [&](const Row& w, auto rc) {
        constexpr int r = decltype(rc)::value;
        const int i = c + 4 * r;
        const int nt = i < M ? min(ftiles4(r), (M - i + 15) >> 4) : 0;
        auto body = [&](auto ntc) {
            constexpr int NT = decltype(ntc)::value;
#pragma unroll
            for (int e = 0; e < 4; ++e)
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    racc[fslot4(r) + t] = mma(w.av[t][e], w.bv[e], racc[fslot4(r) + t]);
        };
        switch (nt) {
            case 1: body(std::integral_constant<int, 1>{}); break;
            case 2: if constexpr (ftiles4(r) >= 2) body(std::integral_constant<int, 2>{}); break;
            case 3: if constexpr (ftiles4(r) >= 3) body(std::integral_constant<int, 3>{}); break;
            case 4: if constexpr (ftiles4(r) >= 4) body(std::integral_constant<int, 4>{}); break;
            default: break;
        }
    };

    __syncthreads();  // the producers' first chunk
    for (int kc = 0; kc < a.nkc; ++kc) {
        const float* Ft = (kc & 1) ? Ft1 : Ft0;
        // row r + 1 is in flight during row r's products
        Row w[2];
        loadrow(w[0], std::integral_constant<int, 0>{}, kc, Ft);
        static_for<0, 16>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            if constexpr (r + 1 < 16) loadrow(w[(r + 1) & 1], std::integral_constant<int, r + 1>{}, kc, Ft);
            rowmma(w[r & 1], rc);
        });
        __syncthreads();
    }

    // c[n, l, m] of column j: instance g0 + j / 2, part j % 2; rows d = 16 t + 4 g + e
    const int ci = g0 + (j >> 1), part = j & 1;
    if (ci >= a.nb) return;
    double* cb = a.coeffs.at(ci);
    static_for<0, 16>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        const int i = c + 4 * r;
        if (i < M) {
            const int l = lbase + 2 * i, nn = M - i;
#pragma unroll
            for (int t = 0; t < ftiles4(r); ++t) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int d = 16 * t + 4 * g + e;
                    if (d < nn) {
                        const int n = l + 2 * d;
                        const double kappa = a.unnormed ? (double)(2 * n + 3) : 1.0;
                        cb[tet_idx(n, l, m) * 2 + part] = (double)racc[fslot4(r) + t][e] * kappa;
                    }
                }
            }
        }
    });
}

__global__ __launch_bounds__(kFT, 1) void fused_fwd_kernel(FusedArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float* Ft0 = reinterpret_cast<float*>(smem_raw);
    float* Ft1 = Ft0 + kFtFwd;
    float* As = Ft1 + kFtFwd;
    const int wi = blockIdx.x / a.ngroups;
    const int g0 = (blockIdx.x - wi * a.ngroups) * kFW;
    const int2 wk = a.work[wi];
    const int m = wk.x, par = wk.y;
    const int M = m < a.nt ? plm_count(a.nt, m, par) : 0;
    if (M <= 0) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // P block: As[i][h] = P_{m+par+2i, m}(z_h), zero padded to 16 NSA rows and 68 columns
    const int K = a.jh, KP = (K + 3) & ~3;
    const int nsa = (M + 15) >> 4;
    const float* P = a.ptab + plm_offset(a.order, m, par, a.jh);
    for (int e = tid; e < 16 * nsa * KP; e += kFT) {
        const int i = e / KP, h = e - i * KP;
        As[i * kFAS + h] = (i < M && h < K) ? P[(long)i * K + h] : 0.f;
    }
    __syncthreads();
    if (wave < 4) {
        switch (nsa) {
            case 1: fused_fwd_produce<1>(a, As, Ft0, Ft1, m, par, g0, wave, lane); break;
            case 2: fused_fwd_produce<2>(a, As, Ft0, Ft1, m, par, g0, wave, lane); break;
            case 3: fused_fwd_produce<3>(a, As, Ft0, Ft1, m, par, g0, wave, lane); break;
            default: fused_fwd_produce<4>(a, As, Ft0, Ft1, m, par, g0, wave, lane); break;
        }
    } else {
        fused_fwd_consume(a, Ft0, Ft1, m, par, M, g0, wave - 4, lane);
    }
}

// ----------------------------------------------------------------- backward
__device__ inline int ftb_index(int i, int inst, int k, int part) {
    return ((i * kFW + inst) * kFK + (k ^ ((2 * inst) & 15))) * 2 + part;
}

// NSH: 16-row strips of the nth/2 half latitudes off the equator
template <int NSH>
__device__ __forceinline__ void fused_bwd_body(const FusedArgs& a, const float* Pb, float* Ft0, float* Ft1, int m,
                                               int par, int M, int g0, int wave, int lane) {
    const int j = lane & 15, g = lane >> 4;
    const int H = a.nth / 2, hoff = a.nth & 1;
    const bool center = hoff && par == 0;
    const long rowlen = 2L * a.ncol;
    const int lbase = m + par;
    const int inst = g0 + wave;  // this wave's instance in the colatitude part
    float* sb = a.spec + (long)(inst < a.nb ? inst : 0) * a.spec_stride + (long)(2 * m + par) * a.jh * rowlen;

    // B operand of the radial part, loaded once: column j = (instance g0 + j / 2,
    // part j % 2), K index d = 16 b + 4 g + s of row i = w + 8 r. Loads are
    // unmasked (index 0 outside the expansion, the value replaced by zero), so
    // that the waits count exactly.
    const int ci = g0 + (j >> 1), part = j & 1;
    const bool cv = ci < a.nb;
    const double* cb = a.coeffs.at(cv ? ci : 0);
    float cB[kFSlots][4];
    static_for<0, kFW>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        const int i = wave + kFW * r;
        const int l = lbase + 2 * i, nn = M - i;
#pragma unroll
        for (int b = 0; b < fblocks(r); ++b)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int d = 16 * b + 4 * g + s;
                const bool ok = cv && i < M && d < nn;
                const int n = l + 2 * d;
                const unsigned idx = ok ? (unsigned)(((n + 1) * (n + 3) * (2 * n + 1)) / 24 + ((l * l) >> 2) + m) * 2 + part : 0u;
                const double x = cb[idx];
                cB[fslot(r) + b][s] = ok ? (float)x : 0.f;
            }
    });

    // A operand of the radial part (rows = nodes j, K = n) of wave-row r, chunk
    // kc: 32-bit element offsets, a uniform row part plus the lane's part (the
    // table holds dp = 64 values of d for every l, so 16 b + 4 g + s < 64 is in
    // it; the row's l is clamped into the table, rows past M form no products)
    const unsigned lstride = (unsigned)a.nkc * kFK * a.dp;  // floats per l in the table
    const unsigned lane_off = (unsigned)j * a.dp + 4 * g;
    auto loadrow = [&](f4 (&av)[4], auto rc, int kc) {
        constexpr int r = decltype(rc)::value;
        const int l = min(lbase + 2 * (wave + kFW * r), a.order - 1);
        const float* rrow = a.rtab + ((unsigned)l * lstride + (unsigned)kc * (kFK * a.dp) + lane_off);
#pragma unroll
        for (int b = 0; b < fblocks(r); ++b) av[b] = *reinterpret_cast<const f4*>(rrow + 16 * b);
    };
    // flm[i][node 4 g + e][column j] of row i = w + 8 r into Ft (two chains of
    // products, even and odd blocks of n); rows i >= M hold zeros
    auto rowmma = [&](const f4 (&av)[4], auto rc, float* Ft) {
        constexpr int r = decltype(rc)::value;
        const int i = wave + kFW * r;
        const int nb = i < M ? min(fblocks(r), (M - i + 15) >> 4) : 0;
        f4 x0 = f4{0, 0, 0, 0}, x1 = f4{0, 0, 0, 0};
        auto body = [&](auto nbc) {
            constexpr int NB = decltype(nbc)::value;
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int bb = 0; bb < NB; ++bb) {
                    if (bb & 1) x1 = mma(av[bb][s], cB[fslot(r) + bb][s], x1);
                    else x0 = mma(av[bb][s], cB[fslot(r) + bb][s], x0);
                }
        };
        switch (nb) {
            case 1: body(std::integral_constant<int, 1>{}); break;
            case 2: if constexpr (fblocks(r) >= 2) body(std::integral_constant<int, 2>{}); break;
            case 3: if constexpr (fblocks(r) >= 3) body(std::integral_constant<int, 3>{}); break;
            case 4: if constexpr (fblocks(r) >= 4) body(std::integral_constant<int, 4>{}); break;
            default: break;
        }
        const f4 x = x0 + x1;
#pragma unroll
        for (int e = 0; e < 4; ++e) Ft[ftb_index(i, j >> 1, 4 * g + e, part)] = x[e];
    };

    const int ksteps = (M + 3) >> 2;
    auto radial = [&](int kc) {
        float* Ft = (kc & 1) ? Ft1 : Ft0;
        f4 av[2][4];
        loadrow(av[0], std::integral_constant<int, 0>{}, kc);
        // row r + 1's loads are in flight during row r's products
        static_for<0, kFW>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            if constexpr (r + 1 < kFW) loadrow(av[(r + 1) & 1], std::integral_constant<int, r + 1>{}, kc);
            rowmma(av[r & 1], rc, Ft);
        });
    };
    // colatitude: rows h = hoff + 16 q + 4 g + e, columns (node j, re/im) of
    // instance w, chunk kc; the LDS operands of step s + 1 are read during the
    // products of step s (rows i < 64 are always in the buffers; rows i >= M
    // hold zeros in both operands)
    const float* Pc = Pb + 16 * NSH * kFAS;
    auto colat = [&](int kc) {
        const float* Ft = (kc & 1) ? Ft1 : Ft0;
        f4 acc[NSH][2];
#pragma unroll
        for (int q = 0; q < NSH; ++q) acc[q][0] = acc[q][1] = f4{0, 0, 0, 0};
        f2 cpart = {0, 0};
        struct Op { f2 bv; float pa[NSH]; float pc; };
        auto ldop = [&](Op& o, int st) {
            const int i = 4 * st + g;
            o.bv = *reinterpret_cast<const f2*>(Ft + ftb_index(i, wave, j, 0));
#pragma unroll
            for (int q = 0; q < NSH; ++q) o.pa[q] = Pb[(16 * q + j) * kFAS + i];
            o.pc = Pc[i];
        };
        Op o;
        ldop(o, 0);
        for (int st = 0; st < ksteps; ++st) {
            Op nx;
            ldop(nx, min(st + 1, 15));
#pragma unroll
            for (int q = 0; q < NSH; ++q) {
                acc[q][0] = mma(o.pa[q], o.bv[0], acc[q][0]);
                acc[q][1] = mma(o.pa[q], o.bv[1], acc[q][1]);
            }
            if (center) {
                cpart[0] += o.pc * o.bv[0];
                cpart[1] += o.pc * o.bv[1];
            }
            o = nx;
        }
        const int k = kc * kFK + j;
        if (inst < a.nb && k < a.ncol) {
            float* o = sb + 2 * k;
#pragma unroll
            for (int q = 0; q < NSH; ++q)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int row = 16 * q + 4 * g + e;
                    if (row < H)
                        *reinterpret_cast<f2*>(o + (hoff + row) * rowlen) = f2{acc[q][0][e], acc[q][1][e]};
                }
        }
        if (center) {
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                float x = cpart[v];
                x += __shfl_xor(x, 16);
                x += __shfl_xor(x, 32);
                cpart[v] = x;
            }
            if (inst < a.nb && k < a.ncol && g == 0) *reinterpret_cast<f2*>(sb + 2 * k) = cpart;
        }
    };

    // chunk kc's flm values into one buffer while chunk kc - 1's colatitude
    // products read the other; one barrier per chunk
    radial(0);
    __syncthreads();
    for (int kc = 1; kc < a.nkc; ++kc) {
        radial(kc);
        colat(kc - 1);
        __syncthreads();
    }
    colat(a.nkc - 1);
}

__global__ __launch_bounds__(kFT, 1) void fused_bwd_kernel(FusedArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float* Ft0 = reinterpret_cast<float*>(smem_raw);
    float* Ft1 = Ft0 + kFtBwd;
    float* Pb = Ft1 + kFtBwd;
    const int wi = blockIdx.x / a.ngroups;
    const int g0 = (blockIdx.x - wi * a.ngroups) * kFW;
    const int2 wk = a.work[wi];
    const int m = wk.x, par = wk.y;
    if (m >= a.nt) return;
    const int M = plm_count(a.nt, m, par);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int H = a.nth / 2, hoff = a.nth & 1;
    const int nsh = (H + 15) >> 4;
    const bool center = hoff && par == 0;
    // Pb[r][i] = P_{m+par+2i, m}(z_{hoff + r}), rows r < 16 nsh, columns i < 64;
    // the equator row (center) at row 16 nsh
    const float* P = a.ptab + plm_offset(a.order, m, par, a.jh);
    const int rows = 16 * nsh + 1;
    for (int e = tid; e < rows * kFRows; e += kFT) {
        const int r = e / kFRows, i = e - r * kFRows;
        float v = 0.f;
        if (i < M) {
            if (r < H) v = P[(long)i * a.jh + hoff + r];
            else if (r == 16 * nsh && center) v = P[(long)i * a.jh];
        }
        Pb[r * kFAS + i] = v;
    }
    __syncthreads();
    switch (nsh) {
        case 1: fused_bwd_body<1>(a, Pb, Ft0, Ft1, m, par, M, g0, wave, lane); break;
        case 2: fused_bwd_body<2>(a, Pb, Ft0, Ft1, m, par, M, g0, wave, lane); break;
        case 3: fused_bwd_body<3>(a, Pb, Ft0, Ft1, m, par, M, g0, wave, lane); break;
        default: fused_bwd_body<4>(a, Pb, Ft0, Ft1, m, par, M, g0, wave, lane); break;
    }
}

// ------------------------------------------------------------------- tables
// fwd: out[((l nkc + kc) dp + d) 16 + kk]; bwd: out[((l nkc + kc) 16 + kk) dp + d];
// R_{l+2d, l}(r_{16 kc + kk}) rounded to FP32, zero outside n < order, k < nr
__global__ void fused_rtab_kernel(const double* rtab, float* out, int order, int nr, int nkc, int dp,
                                  int transpose, long total) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    int d, kk;
    long rest;
    if (transpose) {
        d = (int)(e % dp);
        rest = e / dp;
        kk = (int)(rest % kFK);
        rest /= kFK;
    } else {
        kk = (int)(e % kFK);
        rest = e / kFK;
        d = (int)(rest % dp);
        rest /= dp;
    }
    const int kc = (int)(rest % nkc);
    const int l = (int)(rest / nkc);
    const int n = l + 2 * d, k = kc * kFK + kk;
    out[e] = (n < order && k < nr) ? (float)rtab[rad_idx(n, l) * nr + k] : 0.f;
}

constexpr size_t kFusedLdsFwd = (size_t)(2 * kFtFwd + kFRows * kFAS) * sizeof(float);
constexpr size_t kFusedLdsBwd = (size_t)(2 * kFtBwd + (16 * 4 + 1) * kFAS) * sizeof(float);

// Off by default (ZGL_FUSED=1 turns the fused kernels on): measured at order
// 128 they do not yet beat the stage kernels they replace.
bool fused_env() {
    const char* e = std::getenv("ZGL_FUSED");
    return e && e[0] == '1';
}

int2* upload_pairs(const std::vector<int2>& v, hipStream_t s) {
    if (v.empty()) return nullptr;
    int2* d = dmalloc_n<int2>(v.size(), "work list");
    try {
        check_hip(hipMemcpyAsync(d, v.data(), v.size() * sizeof(int2), hipMemcpyHostToDevice, s), "work list");
    } catch (...) {
        dfree(d);
        throw;
    }
    return d;
}

}  // namespace

bool zt_fused_fits(const OrderTables& t) {
    const Geom& g = t.g;
    return fused_env() && g.order >= 1 && g.order <= 2 * kFRows && g.jh <= 4 * kFSteps && g.nth / 2 <= kFRows &&
           g.ncol > 0;
}

void zt_fused_tables(OrderTables& t, hipStream_t s) {
    if (t.rfuse_f || !zt_fused_fits(t) || !t.rtab || !t.ptab32) return;
    const Geom& g = t.g;
    const int order = g.order;
    t.fuse_nkc = (g.ncol + kFK - 1) / kFK;
    t.fuse_dp = kFRows;  // every d < 64 a kernel may touch is in the table
    const long total = (long)order * t.fuse_nkc * t.fuse_dp * kFK;
    for (int tr = 0; tr < 2; ++tr) {
        float*& out = tr ? t.rfuse_b : t.rfuse_f;
        out = dmalloc_n<float>((size_t)total, "fused radial table");
        fused_rtab_kernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(t.rtab, out, order, g.ncol, t.fuse_nkc,
                                                                         t.fuse_dp, tr, total);
        check_hip(hipGetLastError(), "fused radial table");
    }
    // problems (m, par) by decreasing row count: the costliest workgroups first
    std::vector<int2> fw, bw;
    for (int m = 0; m < order; ++m)
        for (int par = 0; par < 2; ++par) {
            if (plm_count(order, m, par) > 0) fw.push_back(make_int2(m, par));
            bw.push_back(make_int2(m, par));
        }
    auto rows = [&](const int2& p) { return plm_count(order, p.x, p.y); };
    std::stable_sort(fw.begin(), fw.end(), [&](const int2& x, const int2& y) { return rows(x) > rows(y); });
    std::stable_sort(bw.begin(), bw.end(), [&](const int2& x, const int2& y) { return rows(x) > rows(y); });
    t.work_ff = upload_pairs(fw, s); t.n_ff = (int)fw.size();
    t.work_fb = upload_pairs(bw, s); t.n_fb = (int)bw.size();
    check_hip(hipStreamSynchronize(s), "fused tables");
}

static FusedArgs fused_args(const ZtGemmArgs& a, const float* rtab, const int2* work) {
    const OrderTables& t = *a.t;
    const Geom& g = t.g;
    FusedArgs f;
    f.ptab = t.ptab32;
    f.rtab = rtab;
    f.spec = (float*)a.spec;
    f.spec_stride = a.spec_stride * 2;
    f.coeffs = a.coeffs;
    f.work = work;
    f.order = g.order; f.nt = a.nt; f.jh = g.jh; f.ncol = g.ncol; f.nth = g.nth;
    f.nkc = t.fuse_nkc; f.dp = t.fuse_dp; f.unnormed = a.zernike_unnormed; f.nb = a.nb;
    f.ngroups = (a.nb + kFW - 1) / kFW;
    return f;
}

static void allow_fused_lds(const void* kernel, size_t lds) {
    check_hip(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
              "LDS attribute");
}

bool zt_fused_forward(const ZtGemmArgs& a, hipStream_t s) {
    const OrderTables& t = *a.t;
    if (!t.rfuse_f || !zt_fused_fits(t)) return false;
    if (a.nb <= 0 || a.nt <= 0 || t.n_ff == 0) return true;
    const FusedArgs f = fused_args(a, t.rfuse_f, t.work_ff);
    allow_fused_lds(reinterpret_cast<const void*>(&fused_fwd_kernel), kFusedLdsFwd);
    fused_fwd_kernel<<<(unsigned)t.n_ff * (unsigned)f.ngroups, kFT, kFusedLdsFwd, s>>>(f);
    check_hip(hipGetLastError(), "fused forward launch");
    return true;
}

bool zt_fused_backward(const ZtGemmArgs& a, hipStream_t s) {
    const OrderTables& t = *a.t;
    if (!t.rfuse_b || !zt_fused_fits(t)) return false;
    if (a.nb <= 0 || a.nt <= 0 || t.n_fb == 0) return true;
    const FusedArgs f = fused_args(a, t.rfuse_b, t.work_fb);
    allow_fused_lds(reinterpret_cast<const void*>(&fused_bwd_kernel), kFusedLdsBwd);
    fused_bwd_kernel<<<(unsigned)t.n_fb * (unsigned)f.ngroups, kFT, kFusedLdsBwd, s>>>(f);
    check_hip(hipGetLastError(), "fused backward launch");
    return true;
}

}  // namespace zgl
