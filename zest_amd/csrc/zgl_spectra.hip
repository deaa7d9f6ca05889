// Power spectra of coefficient arrays: st::power_spectrum,
// st::cross_power_spectrum and zt::power_spectrum
// (include/power_spectra.hpp:43-141), batched.
//
// One thread per output entry (l for SH, (n, l) for Zernike) sums its m run
// of {cos, sin} pairs in zest's order (m = 0 cos part first, then m = 1..l);
// the pairs of one entry are contiguous, and the entries of one instance are
// contiguous runs in the same order, so a wave reads one span of the array.
#include <hip/hip_runtime.h>

#include <cmath>

#include "zgl_internal.hpp"

namespace zgl {

namespace {

__global__ void spectrum_kernel(int zernike, int order, long nent, Inst<const double> a,
                                Inst<const double> b, double norm, double* out, long os) {
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= nent) return;
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
    double s = x[0] * y[0];
    for (int m = 1; m <= l; ++m) s += x[2 * m] * y[2 * m] + x[2 * m + 1] * y[2 * m + 1];
    out[inst * os + e] = norm * s;
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
    spectrum_kernel<<<dim3((unsigned)((nent + 255) / 256), nb), 256, 0, s>>>(
        kind == ZGL_ZERNIKE ? 1 : 0, order, nent, a, b, norm, out, out_stride);
    check_hip(hipGetLastError(), "power spectrum launch");
}

}  // namespace zgl
