// Power spectra of coefficient arrays: st::power_spectrum,
// st::cross_power_spectrum and zt::power_spectrum
// (include/power_spectra.hpp:43-141), batched.
//
// Sixteen lanes per output entry (l for SH, (n, l) for Zernike): the {cos, sin}
// pairs of one entry are contiguous, so the group reads 16 consecutive pairs
// (256 bytes) per step, each lane summing m = k, k + 16, ... (the m = 0 sin
// part excluded, as zest excludes it), and the group reduces with shuffles.
#include <hip/hip_runtime.h>

#include <cmath>

#include "zgl_internal.hpp"

namespace zgl {

namespace {

constexpr int kGroup = 8;                // lanes per entry
constexpr int kSpecThreads = 256;
constexpr int kEntriesPerBlock = kSpecThreads / kGroup;

// VEC: every instance base is 16-byte aligned, so a (re, im) pair is one
// 16-byte load; otherwise two 8-byte loads (instance strides are counted in
// doubles and need not keep pairs aligned).
template <bool CROSS, bool VEC>
__global__ __launch_bounds__(kSpecThreads) void spectrum_kernel(int zernike, long nent, Inst<const double> a,
                                                                Inst<const double> b, double norm, double* out,
                                                                long os) {
    const long e = blockIdx.x * (long)kEntriesPerBlock + (threadIdx.x / kGroup);
    const int k = threadIdx.x % kGroup;
    if (e >= nent) return;  // whole groups leave together
    const int inst = blockIdx.y;
    long first;  // pair index of m = 0
    int l;
    if (zernike) {
        // e = rad_idx(n, l) = ((n+1)^2 >> 2) + (l >> 1)
        int n = max(0, (int)(2.0 * sqrt((double)e)) - 2);
        while (((long)(n + 2) * (n + 2) >> 2) <= e) ++n;
        l = 2 * (int)(e - (((long)(n + 1) * (n + 1)) >> 2)) + (n & 1);
        first = tet_idx(n, l, 0);
    } else {
        l = (int)e;
        first = tri_idx(l, 0);
    }
    const double* x = a.at(inst) + 2 * first;
    const double* y = b.at(inst) + 2 * first;
    double s = 0;
    if constexpr (VEC) {
        const double2* x2 = reinterpret_cast<const double2*>(x);
        const double2* y2 = reinterpret_cast<const double2*>(y);
#pragma unroll 4
        for (int m = k; m <= l; m += kGroup) {
            const double2 u = x2[m];
            const double2 v = CROSS ? y2[m] : u;
            s += u.x * v.x;
            if (m > 0) s += u.y * v.y;
        }
    } else {
#pragma unroll 4
        for (int m = k; m <= l; m += kGroup) {
            const double ur = x[2 * m], ui = x[2 * m + 1];
            const double vr = CROSS ? y[2 * m] : ur, vi = CROSS ? y[2 * m + 1] : ui;
            s += ur * vr;
            if (m > 0) s += ui * vi;
        }
    }
#pragma unroll
    for (int d = kGroup / 2; d > 0; d >>= 1) s += __shfl_xor(s, d, kGroup);
    if (k == 0) out[inst * os + e] = norm * s;
}

}  // namespace

long spectrum_size(zgl_kind kind, int order) {
    return kind == ZGL_ZERNIKE ? (long)(order + 1) * (order + 1) / 4 : order;
}

void power_spectrum(zgl_kind kind, int sh_norm, int order, Inst<const double> a,
                    Inst<const double> b, double* out, long out_stride, int nb, hipStream_t s) {
    const long nent = spectrum_size(kind, order);
    if (nent == 0 || nb == 0) return;
    const double norm = sh_norm == ZGL_SHNORM_QM ? 1.0 : 1.0 / (4.0 * M_PI);
    const dim3 grid((unsigned)((nent + kEntriesPerBlock - 1) / kEntriesPerBlock), nb);
    const bool cross = a.base != b.base || a.stride != b.stride || a.offs != b.offs;
    auto aligned = [](const Inst<const double>& v) {
        return !v.offs && reinterpret_cast<uintptr_t>(v.base) % 16 == 0 && v.stride % 2 == 0;
    };
    const bool vec = aligned(a) && aligned(b);
    const int zk = kind == ZGL_ZERNIKE ? 1 : 0;
    if (cross && vec)
        spectrum_kernel<true, true><<<grid, kSpecThreads, 0, s>>>(zk, nent, a, b, norm, out, out_stride);
    else if (cross)
        spectrum_kernel<true, false><<<grid, kSpecThreads, 0, s>>>(zk, nent, a, b, norm, out, out_stride);
    else if (vec)
        spectrum_kernel<false, true><<<grid, kSpecThreads, 0, s>>>(zk, nent, a, b, norm, out, out_stride);
    else
        spectrum_kernel<false, false><<<grid, kSpecThreads, 0, s>>>(zk, nent, a, b, norm, out, out_stride);
    check_hip(hipGetLastError(), "power spectrum launch");
}

}  // namespace zgl
