// P_lm(z) for one (m, z): the column recursion in l of
// st::PlmRecursion::plm_impl (include/plm_recursion.hpp:229-354) with its
// coefficients (src/plm_recursion.cpp:181-247), the 1e-280 seed scaling that
// keeps high-m seeds from underflowing (u^m applied lazily through the factor
// us), and the real normalisation sqrt(2) for m > 0. Emits emit(l, P_lm(z))
// for l = m .. order-1 in increasing l. Shared by the table kernel and the
// Legendre kernels that recompute their P block instead of reading the table.
#pragma once

#include <hip/hip_runtime.h>

namespace zgl {

__device__ inline double plm_sqrt(int k) { return sqrt((double)k); }

template <class Emit>
__device__ inline void plm_column(int order, int m, double z, int qm, double phase, Emit&& emit) {
    const double inv_sqrt_4pi = 0.5 / sqrt(M_PI);
    if (m >= order) return;
    if (m == 0) {
        double p2 = qm ? inv_sqrt_4pi : 1.0;
        emit(0, p2);
        if (order == 1) return;
        double p1 = qm ? z * (plm_sqrt(3) * inv_sqrt_4pi) : z * plm_sqrt(3);
        emit(1, p1);
        for (int l = 2; l < order; ++l) {
            const double a = plm_sqrt(2 * l - 1) * plm_sqrt(2 * l + 1) / (double)l;
            const double b = ((double)l - 1.0) * plm_sqrt(2 * l + 1) / (plm_sqrt(2 * l - 3) * (double)l);
            const double p = a * z * p1 - b * p2;
            emit(l, p);
            p2 = p1;
            p1 = p;
        }
        return;
    }
    const double u = sqrt((1.0 - z) * (1.0 + z));
    double pmm = qm ? 1.0e-280 * sqrt(2.0) * inv_sqrt_4pi : 1.0e-280 * sqrt(2.0);
    double us = 1.0 / 1.0e-280;
    const int msteps = m < order - 1 ? m : order - 2;
    for (int i = 1; i <= msteps; ++i) {
        us *= u;
        pmm *= phase * plm_sqrt(2 * i + 1) / plm_sqrt(2 * i);
    }
    if (m == order - 1) {
        us *= u;
        emit(m, us * (phase * pmm * plm_sqrt(2 * order - 1) / plm_sqrt(2 * order - 2)));
        return;
    }
    double p2 = pmm;
    double p1 = z * (plm_sqrt(2 * m + 3) * pmm);
    emit(m, p2 * us);
    emit(m + 1, p1 * us);
    for (int l = m + 2; l < order; ++l) {
        const double a = plm_sqrt(2 * l - 1) * plm_sqrt(2 * l + 1) / (plm_sqrt(l - m) * plm_sqrt(l + m));
        const double b = plm_sqrt(2 * l + 1) * plm_sqrt(l + m - 1) * plm_sqrt(l - m - 1) /
                         (plm_sqrt(l - m) * plm_sqrt(l + m) * plm_sqrt(2 * l - 3));
        const double p = z * a * p1 - b * p2;
        emit(l, p * us);
        p2 = p1;
        p1 = p;
    }
}

}  // namespace zgl
