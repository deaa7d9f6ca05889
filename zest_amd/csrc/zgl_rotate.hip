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

__global__ __launch_bounds__(256, 2) void rotate_kernel(RotArgs a) {
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
    const bool wave_active = 2 * (i0 + 16 * wave) + p <= l;  // wave-uniform

    // pair index of a column's m' = 0 element, and its instance
    auto locate = [&](int col, int& inst) -> long {
        inst = col;
        if (a.zt) {
            inst = col / cnt;
            return tet_idx(l + 2 * (col - inst * cnt), l, 0);
        }
        return tri_idx(l, 0);
    };

    // Staging: thread tid fills column tid % 64 of the chunk at m' = k0 +
    // tid / 64 + 4u, u < 16; its column and instance are fixed.
    const int sc = tid & (kRotCols - 1), sk = tid / kRotCols;
    const int scol = c0 + sc;
    const bool scolok = scol < ncols;
    const double* sp = nullptr;
    d2 step = {1.0, 0.0};  // e^{-4 i gamma'} (pass 0)
    double sga = 0.0;
    if (scolok) {
        int inst;
        const long base = locate(scol, inst);
        sp = a.src.at(inst) + 2 * base;
        double al, be;
        converted(a, inst, al, be, sga);
        double s4, c4;
        sincos(-4.0 * sga, &s4, &c4);
        step = d2{c4, s4};
    }

    d4 are[4], aim[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) are[c] = aim[c] = d4{0, 0, 0, 0};

    for (int k0 = 0; k0 <= l; k0 += kRotK) {
        // the d values of this chunk's MFMA steps (independent of the LDS)
        double xa[kRotK / 8], xb[kRotK / 8];
#pragma unroll
        for (int s = 0; s < kRotK / 8; ++s) {
            const int kk = 4 * s + g;
            const int mre = k0 + 2 * kk + pa, mim = k0 + 2 * kk + pb;
            xa[s] = xb[s] = 0.0;
            if (wave_active && rowok && mre <= l)
                xa[s] = (mre ? 2.0 : 1.0) *
                        (a.stage == 0 ? dl[(long)(l + 1) * mrow + mre] : dl[(long)(l + 1) * mre + mrow]);
            if (wave_active && rowok && mim <= l)
                xb[s] = 2.0 * (a.stage == 0 ? dl[(long)(l + 1) * mrow + mim] : dl[(long)(l + 1) * mim + mrow]);
        }
        __syncthreads();  // the previous chunk's LDS reads are done
        // two halves of 8 independent loads each
        double s0 = 0.0, c0v = 1.0;
        if (a.stage == 0 && scolok) sincos(-(double)(k0 + sk) * sga, &s0, &c0v);
        d2 ph = {c0v, s0};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            d2 v[kRotK / 8];
#pragma unroll
            for (int u = 0; u < kRotK / 8; ++u) {
                const int mp = k0 + sk + 4 * (8 * h + u);
                v[u] = (scolok && mp <= l) ? *reinterpret_cast<const d2*>(sp + 2 * mp) : d2{0.0, 0.0};
            }
            if (a.stage == 0) {
                // v_m' scaled to the complex form, times e^{-i m' gamma'}
#pragma unroll
                for (int u = 0; u < kRotK / 8; ++u) {
                    const int mp = k0 + sk + 4 * (8 * h + u);
                    const d2 x = mp ? d2{v[u][0] * kInvSqrt2, v[u][1] * -kInvSqrt2} : v[u];
                    v[u] = d2{x[0] * ph[0] - x[1] * ph[1], x[0] * ph[1] + x[1] * ph[0]};
                    ph = d2{ph[0] * step[0] - ph[1] * step[1], ph[0] * step[1] + ph[1] * step[0]};
                }
            }
#pragma unroll
            for (int u = 0; u < kRotK / 8; ++u) Xs[(sk + 4 * (8 * h + u)) * kRotCols + sc] = v[u];
        }
        __syncthreads();
        if (wave_active) {
#pragma unroll
            for (int s = 0; s < kRotK / 8; ++s) {
                const int kk = 4 * s + g;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const double bre = Xs[(2 * kk + pa) * kRotCols + 16 * c + j][0];
                    const double bim = Xs[(2 * kk + pb) * kRotCols + 16 * c + j][1];
                    are[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[s], bre, are[c], 0, 0, 0);
                    aim[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(xb[s], bim, aim[c], 0, 0, 0);
                }
            }
        }
    }
    if (!wave_active) return;

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

// ---------------------------------------------------------------- fused
// Degrees l <= 255: one workgroup holds every m of CT columns in LDS
// ((l+1) CT <= 4096 complex values) and runs the whole rotation there:
// stage (scale, gamma phase) -> X(+90) into registers -> beta phase back into
// LDS -> X(-90) -> alpha phase, rescale, store. The coefficients cross HBM
// once each way. The MFMA tiles are (row parity, 16-row block, 16-column
// block) tasks, at most four per wave, advanced in lockstep so that the d
// values of the next step (read from L2) are in flight during the current
// step's MFMAs.
constexpr int kFusedMaxL = 255;
constexpr int kFusedElems = 4096;

__host__ __device__ inline int fused_cols(int l) { return l < 64 ? 64 : (l < 128 ? 32 : 16); }

// LDS: the (l+1) x CT complex block, then per column its coefficient block
// offset and the converted angles
constexpr size_t kFusedLds = kFusedElems * sizeof(d2) + 64 * (sizeof(long) + 3 * sizeof(double));

__global__ __launch_bounds__(256, 2) void rotate_fused_kernel(RotArgs a) {
    extern __shared__ __align__(16) unsigned char fused_smem[];
    d2* Xs = reinterpret_cast<d2*>(fused_smem);
    long* coff = reinterpret_cast<long*>(Xs + kFusedElems);     // [CT] doubles from coefficient base
    double* cang = reinterpret_cast<double*>(coff + 64);        // [CT][3] alpha', beta', gamma'
    const int4 w = a.work[blockIdx.x];
    const int l = w.x;
    const int CT = fused_cols(l), CB = CT / 16;
    const int c0 = w.y * CT;
    const int cnt = a.zt ? (a.order - l + 1) / 2 : 1;
    const int ncols = a.nb * cnt;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int j = lane & 15, g = lane >> 4;
    const double* dl = a.dtab + wig_base(l);
    const int L1 = l + 1;
    const int nel = L1 * CT;

    // ---- per column: where its block starts (a.src and a.dst share the layout), its angles
    if (tid < CT) {
        const int col = c0 + tid;
        long off = -1;
        double al = 0.0, be = 0.0, ga = 0.0;
        if (col < ncols) {
            int inst = col;
            long base = tri_idx(l, 0);
            if (a.zt) {
                inst = col / cnt;
                base = tet_idx(l + 2 * (col - inst * cnt), l, 0);
            }
            off = (a.src.at(inst) - a.src.base) + 2 * base;
            converted(a, inst, al, be, ga);
        }
        coff[tid] = off;
        cang[3 * tid] = al;
        cang[3 * tid + 1] = be;
        cang[3 * tid + 2] = ga;
    }
    __syncthreads();

    // ---- stage: element e = (column e / L1, m' e % L1): consecutive lanes read
    // consecutive pairs of one block. A thread's elements step by 256, so its
    // (column, m') pair is stepped by (256 / L1, 256 % L1) with one carry.
    const int cq = 256 / L1, cr = 256 - cq * L1;
    auto advance = [&](int& c, int& mp) {
        c += cq;
        mp += cr;
        if (mp >= L1) {
            mp -= L1;
            ++c;
        }
    };
    {
        int c = tid / L1, mp = tid - (tid / L1) * L1;
        for (int e0 = 0; e0 < nel; e0 += 256 * 8) {
            d2 v[8];
            int cu[8], mu[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + 256 * u + tid;
                cu[u] = c;
                mu[u] = mp;
                v[u] = d2{0.0, 0.0};
                if (e < nel) {
                    const long off = coff[c];
                    if (off >= 0) v[u] = *reinterpret_cast<const d2*>(a.src.base + off + 2 * mp);
                }
                advance(c, mp);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + 256 * u + tid;
                if (e < nel) Xs[mu[u] * CT + cu[u]] = v[u];
            }
        }
    }
    // In place on Xs: thread tid takes column tid % CT at rows tid / CT +
    // (256 / CT) u and multiplies by e^{-i m x_c}, x_c angle `which` of its
    // column, the power stepped by a complex product (two sincos per thread).
    // mode 0: complex form first (m > 0 pairs times (1/sqrt2, -1/sqrt2));
    // mode 2: back to {cos, sin} pairs after (times (sqrt2, -sqrt2)).
    auto phase_pass = [&](int which, int mode) {
        const int c = tid % CT, r0 = tid / CT, stride = 256 / CT;
        const double x = cang[3 * c + which];
        double ss, cs;
        sincos(-(double)stride * x, &ss, &cs);
        const d2 step = {cs, ss};
        sincos(-(double)r0 * x, &ss, &cs);
        d2 ph = {cs, ss};
        for (int m = r0; m <= l; m += stride) {
            d2 v = Xs[m * CT + c];
            if (mode == 0 && m) v = d2{v[0] * kInvSqrt2, v[1] * -kInvSqrt2};
            v = d2{v[0] * ph[0] - v[1] * ph[1], v[0] * ph[1] + v[1] * ph[0]};
            if (mode == 2 && m) v = d2{v[0] * kSqrt2, v[1] * -kSqrt2};
            Xs[m * CT + c] = v;
            ph = d2{ph[0] * step[0] - ph[1] * step[1], ph[0] * step[1] + ph[1] * step[0]};
        }
    };
    __syncthreads();
    phase_pass(2, 0);  // gamma'

    // ---- tasks of this wave: row tasks (row parity, 16-row block) w, w + 4,
    // .., each with all CB column blocks; four (row task, column block) slots
    // per wave, slot k = (row task k / CB, column block k % CB)
    const int rb0 = (l / 2 + 1 + 15) / 16;        // parity 0: rows m = 0, 2, .., <= l
    const int rb1 = ((l - 1) / 2 + 1 + 15) / 16;  // parity 1 (l >= 1)
    const int NR = rb0 + rb1;                     // row tasks
    const int nrt = 4 / CB;                       // row tasks per wave
    int rp[4], rrow[4];
    bool rok[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int t = wave + 4 * r;
        rok[r] = r < nrt && t < NR;
        rp[r] = t < rb0 ? 0 : 1;
        rrow[r] = 2 * (16 * (t - (rp[r] ? rb0 : 0))) + rp[r];  // m of row 0; row i is rrow + 2i
    }
    int tr[4], tcol[4];
    bool tok[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        tr[k] = k / CB;
        tcol[k] = 16 * (k % CB);
        tok[k] = rok[tr[k]];
    }
    int tp[4], trow[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        tp[k] = rp[tr[k]];
        trow[k] = rrow[tr[k]];
    }
    // steps of 4 values of m' of one parity: the longer of the two parity runs
    const int steps = (l / 2 + 1 + 3) / 4;

    d4 are[4], aim[4];
    // Per lane and row task: pointers to the d values of step 0 (the row m of
    // the lane, columns m' = 2g + parity and its opposite parity), advanced by
    // 8 values of m' per step, and the last step each is inside the degree;
    // per slot the same for the B operand in LDS. No index arithmetic per step.
    const int ld = l + 1;
    const int bstep = 8 * CT;
    const d2 *bre_p[4], *bim_p[4];
    int blre[4], blim[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int pa = (l + tp[k]) & 1;
        const int cre = 2 * g + pa, cim = 2 * g + (pa ^ 1);
        bre_p[k] = Xs + cre * CT + tcol[k] + j;
        bim_p[k] = Xs + cim * CT + tcol[k] + j;
        blre[k] = cre <= l ? (l - cre) >> 3 : -1;
        blim[k] = cim <= l ? (l - cim) >> 3 : -1;
    }
    auto product = [&](int pass) {
#pragma unroll
        for (int k = 0; k < 4; ++k) are[k] = aim[k] = d4{0, 0, 0, 0};
        const long dstep = pass == 0 ? 8 : 8L * ld;
        const double *pre[4], *pim[4];
        int lre[4], lim[4];
        bool z0[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int pa = (l + rp[r]) & 1;
            const int m = rrow[r] + 2 * j;
            const bool mok = rok[r] && m <= l;
            const int cre = 2 * g + pa, cim = 2 * g + (pa ^ 1);
            pre[r] = pass == 0 ? dl + (long)ld * m + cre : dl + (long)ld * cre + m;
            pim[r] = pass == 0 ? dl + (long)ld * m + cim : dl + (long)ld * cim + m;
            lre[r] = mok && cre <= l ? (l - cre) >> 3 : -1;
            lim[r] = mok && cim <= l ? (l - cim) >> 3 : -1;
            z0[r] = cre == 0;  // m' = 0 at step 0: the weight 1 instead of 2
        }
        // d values of step s for each row task
        auto load = [&](int s, double (&xa)[4], double (&xb)[4]) {
            const long o = s * dstep;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                xa[r] = s <= lre[r] ? ((s == 0 && z0[r]) ? 1.0 : 2.0) * pre[r][o] : 0.0;
                xb[r] = s <= lim[r] ? 2.0 * pim[r][o] : 0.0;
            }
        };
        // two steps in flight ahead of the MFMAs
        double xa0[4], xb0[4], xa1[4], xb1[4];
        load(0, xa0, xb0);
        load(1, xa1, xb1);
        for (int s = 0; s < steps; ++s) {
            double na[4] = {0, 0, 0, 0}, nb[4] = {0, 0, 0, 0};
            if (s + 2 < steps) load(s + 2, na, nb);
            const int bo = s * bstep;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (!tok[k]) continue;
                const double bre = s <= blre[k] ? bre_p[k][bo][0] : 0.0;
                const double bim = s <= blim[k] ? bim_p[k][bo][1] : 0.0;
                are[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa0[tr[k]], bre, are[k], 0, 0, 0);
                aim[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(xb0[tr[k]], bim, aim[k], 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                xa0[r] = xa1[r];
                xb0[r] = xb1[r];
                xa1[r] = na[r];
                xb1[r] = nb[r];
            }
        }
    };

    __syncthreads();
    product(0);
    __syncthreads();  // every X(+90) read of Xs is done
    // y back into Xs, times e^{-i m beta'}
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!tok[k]) continue;
        const int c = tcol[k] + j;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = trow[k] + 2 * (g + 4 * r);
            if (m <= l) Xs[m * CT + c] = d2{are[k][r], aim[k][r]};
        }
    }
    __syncthreads();
    phase_pass(1, 1);
    __syncthreads();
    product(1);
    __syncthreads();  // every X(-90) read of Xs is done
    // z into Xs, times e^{-i m alpha'} as {cos, sin} pairs, then out with the
    // staging's mapping
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!tok[k]) continue;
        const int c = tcol[k] + j;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = trow[k] + 2 * (g + 4 * r);
            if (m <= l) Xs[m * CT + c] = d2{are[k][r], aim[k][r]};
        }
    }
    __syncthreads();
    phase_pass(0, 2);
    __syncthreads();
    {
        int c = tid / L1, mp = tid - (tid / L1) * L1;
        for (int e = tid; e < nel; e += 256) {
            const long off = coff[c];
            if (off >= 0) *reinterpret_cast<d2*>(a.dst.base + off + 2 * mp) = Xs[mp * CT + c];
            advance(c, mp);
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

std::vector<int4> rotation_fused_work(bool zt, int order, int nb) {
    std::vector<int4> w;
    for (int l = 1; l < order && l <= kFusedMaxL; ++l) {
        const int cnt = zt ? (order - l + 1) / 2 : 1;
        const long ncols = (long)nb * cnt;
        const int CT = fused_cols(l);
        for (int c = 0; c < (int)((ncols + CT - 1) / CT); ++c) w.push_back(make_int4(l, c, 0, 0));
    }
    return w;
}

std::vector<int4> rotation_work(bool zt, int order, int nb) {
    std::vector<int4> w;
    for (int l = kFusedMaxL + 1; l < order; ++l) {
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

void rotate_fused(bool zt, int order, const double* dtab, const int4* work, int nwork, Inst<double> coeffs,
                  const double* angles, int coordinate, int nb, hipStream_t s) {
    if (nwork == 0 || nb == 0) return;
    RotArgs a{dtab, work, zt ? 1 : 0, order, nb, Inst<const double>(coeffs), coeffs, angles, coordinate, 0};
    check_hip(hipFuncSetAttribute(reinterpret_cast<const void*>(&rotate_fused_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFusedLds),
              "LDS attribute");
    rotate_fused_kernel<<<nwork, 256, kFusedLds, s>>>(a);
    check_hip(hipGetLastError(), "rotation launch");
}

void rotate_pass(int pass, bool zt, int order, const double* dtab, const int4* work, int nwork,
                 Inst<double> coeffs, double* temp, long temp_stride, const double* angles, int coordinate,
                 int nb, hipStream_t s) {
    if (nwork == 0 || nb == 0) return;
    const Inst<double> tmp{temp, temp_stride, nullptr};
    RotArgs a{dtab, work, zt ? 1 : 0, order, nb,
              pass == 0 ? Inst<const double>(coeffs) : Inst<const double>(tmp),
              pass == 0 ? tmp : coeffs, angles, coordinate, pass};
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
