// Device generation of the FP64 tables of one order: Gauss-Legendre nodes and
// weights, the associated Legendre table P_lm(z_j) and the radial Zernike
// table R_nl(r_k). All tables are built once per (plan, order) and cached.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "zgl_internal.hpp"
#include "zgl_plm.hpp"

namespace zgl {

namespace {

// Legendre P_n and P_n' by the three-term recurrence.
__device__ void legendre_p_dp(int n, double x, double& p, double& dp) {
    double p0 = 1.0, p1 = x;
    if (n == 0) { p = 1.0; dp = 0.0; return; }
    for (int k = 1; k < n; ++k) {
        double p2 = ((double)(2 * k + 1) * x * p1 - (double)k * p0) / (double)(k + 1);
        p0 = p1;
        p1 = p2;
    }
    p = p1;
    dp = (double)n * (x * p1 - p0) / (x * x - 1.0);
}

// Double-double arithmetic (error-free sums, FMA products) for the final
// Newton steps and the weights.
struct dd { double hi, lo; };
__device__ inline dd dd_two_sum(double a, double b) {
    const double s = a + b, bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ inline dd dd_fast_two_sum(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
__device__ inline dd dd_add(dd x, dd y) {
    const dd s = dd_two_sum(x.hi, y.hi);
    return dd_fast_two_sum(s.hi, s.lo + x.lo + y.lo);
}
__device__ inline dd dd_neg(dd x) { return {-x.hi, -x.lo}; }
__device__ inline dd dd_mul(dd x, dd y) {
    const double p = x.hi * y.hi, e = fma(x.hi, y.hi, -p);
    return dd_fast_two_sum(p, e + (x.hi * y.lo + x.lo * y.hi));
}
__device__ inline dd dd_mul_d(dd x, double y) {
    const double p = x.hi * y, e = fma(x.hi, y, -p);
    return dd_fast_two_sum(p, e + x.lo * y);
}
__device__ inline dd dd_div(dd x, dd y) {
    const double q1 = x.hi / y.hi;
    dd r = dd_add(x, dd_neg(dd_mul_d(y, q1)));
    const double q2 = r.hi / y.hi;
    r = dd_add(r, dd_neg(dd_mul_d(y, q2)));
    const double q3 = r.hi / y.hi;
    return dd_add(dd_fast_two_sum(q1, q2), dd{q3, 0.0});
}

__device__ void legendre_p_dp_dd(int n, dd x, dd& p, dd& dp) {
    dd p0{1.0, 0.0}, p1 = x;
    for (int k = 1; k < n; ++k) {
        const dd t = dd_add(dd_mul_d(dd_mul(x, p1), (double)(2 * k + 1)), dd_neg(dd_mul_d(p0, (double)k)));
        p0 = p1;
        p1 = dd_div(t, dd{(double)(k + 1), 0.0});
    }
    p = p1;
    const dd num = dd_mul_d(dd_add(dd_mul(x, p1), dd_neg(p0)), (double)n);
    dp = dd_div(num, dd_add(dd_mul(x, x), dd{-1.0, 0.0}));
}

// Gauss-Legendre nodes (ascending, cos style) and weights of n points, one
// thread per node pair: Newton's method on P_n from Tricomi's initial guess,
// then two Newton steps and the weight 2 / ((1 - x^2) P_n'(x)^2) in
// double-double, so both round to within about an ulp. Same node set as
// gl::gl_nodes_and_weights<UnpackedLayout, cos> (include/gauss_legendre.hpp:1549-1562),
// which uses tables below 70 nodes and Bogaert's formulas above.
__global__ void gl_kernel(int n, double* nodes, double* weights) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (n + 1) / 2) return;
    if (n == 1) {
        nodes[0] = 0.0;
        weights[0] = 2.0;
        return;
    }
    double x = cos(M_PI * ((double)k + 0.75) / ((double)n + 0.5));
    double p, dp;
    for (int it = 0; it < 100; ++it) {
        legendre_p_dp(n, x, p, dp);
        const double dx = p / dp;
        x -= dx;
        if (fabs(dx) <= 1e-15) break;
    }
    dd xd{x, 0.0}, pd, dpd;
    if ((n & 1) && k == n / 2) {
        xd.hi = 0.0;
    } else {
        for (int it = 0; it < 2; ++it) {
            legendre_p_dp_dd(n, xd, pd, dpd);
            xd = dd_add(xd, dd_neg(dd_div(pd, dpd)));
        }
    }
    legendre_p_dp_dd(n, xd, pd, dpd);
    const dd w = dd_div(dd{2.0, 0.0}, dd_mul(dd_add(dd{1.0, 0.0}, dd_neg(dd_mul(xd, xd))), dd_mul(dpd, dpd)));
    nodes[n - 1 - k] = xd.hi;
    nodes[k] = -xd.hi;
    weights[n - 1 - k] = w.hi;
    weights[k] = w.hi;
}

// North-half weights w[nth/2 + h] and nodes z[nth/2 + h]; for Zernike also
// the radial weights r_k^2 v_k with r_k = (1 + x_k) / 2
// (include/zernike_glq_transformer.hpp:538-546, 741-746).
__global__ void half_nodes_kernel(int nth, const double* z, const double* w, double* zh,
                                  double* wh) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    const int jh = (nth + 1) / 2;
    if (h >= jh) return;
    zh[h] = z[nth / 2 + h];
    wh[h] = w[nth / 2 + h];
}

__global__ void radial_nodes_kernel(int nr, double* x, const double* v, double* wr) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr) return;
    const double r = 0.5 * (1.0 + x[k]);
    x[k] = r;
    wr[k] = r * r * v[k];
}

// P_lm(z_h) for one (h, m) (plm_column, zgl_plm.hpp), in the parity-split
// layout [m][par][l'][h], l = m + par + 2 l'.
__global__ void plm_kernel(int order, int jh, int qm, double phase, const double* zh,
                           double* ptab) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int h = gid % jh;
    const int m = gid / jh;
    if (m >= order) return;
    plm_column(order, m, zh[h], qm, phase, [&](int l, double v) {
        const int par = (l - m) & 1;
        const int lp = (l - m - par) >> 1;
        ptab[plm_offset(order, m, par, jh) + (long)lp * jh + h] = v;
    });
}

// R_nl(r_k): the recursion of zt::RadialZernikeRecursion::zernike
// (include/radial_zernike_recursion.hpp:151-288) with the coefficients of
// src/radial_zernike_recursion.cpp:91-116, one thread per (radius, l): the
// sequence R_ll, R_{l+2,l}, R_{l+4,l}, ... of one l depends on other l only
// through its seeds R_ll = r^l (formed by the same repeated products as the
// recursion's diagonal R_nn = r R_{n-1,n-1}) and R_{l+2,l+2}. Every recursion
// input is unnormalised, so the normed table is the unnormed one times
// sqrt(2n+3), applied as each value is stored.
__global__ void radial_kernel(int order, int nr, int normed, const double* rk, double* rtab) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    if (k >= nr || l >= order) return;
    const double r = rk[k];
    const double r2 = r * r;
    auto store = [&](int n, double v) {
        rtab[rad_idx(n, l) * nr + k] = normed ? v * sqrt(2.0 * n + 3.0) : v;
    };
    // powers r^l and r^(l+2) as the diagonal recursion forms them
    double dl = 1.0;
    for (int i = 1; i <= l; ++i) dl = i == 1 ? r : (i == 2 ? r2 : dl * r);
    store(l, dl);
    if (l + 2 >= order) return;
    const int n2 = l + 2;
    const double dl2 = n2 == 2 ? r2 : dl * r * r;
    double p2;  // R_{l+2,l}
    if (n2 == 2) p2 = 2.5 * r2 - 1.5;
    else if (n2 == 3) p2 = (3.5 * r2 - 2.5) * r;
    else p2 = (n2 + 0.5) * dl2 - (n2 - 0.5) * dl;
    store(n2, p2);
    double p4 = dl;  // R_{n-4,l}
    for (int n = l + 4; n < order; n += 2) {
        const double dn = (double)n, dL = (double)l;
        const double c0 = (dn - dL) * (dn + dL + 1.0) * (2.0 * dn - 3.0);
        const double c1 = (2.0 * dn - 1.0) * (2.0 * dn + 1) * (2.0 * dn - 3.0);
        const double c2 = -0.5 * (2.0 * dn - 1.0) *
                          ((2 * dL + 1.0) * (2.0 * dL + 1.0) + (2.0 * dn + 1) * (2.0 * dn - 3.0));
        const double c3 = -(dn - dL - 2.0) * (dn + dL - 1.0) * (2.0 * dn + 1);
        const double inv = 1.0 / c0;
        const double k1 = c1 * inv, k2 = c2 * inv, k3 = c3 * inv;
        const double v = (k2 + k1 * r2) * p2 + k3 * p4;
        store(n, v);
        p4 = p2;
        p2 = v;
    }
}

__global__ void to_fp32_kernel(const double* in, float* out, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

template <class G>
__global__ void residual_kernel(Inst<const G> g, double* r, long rs, long n) {
    const G* gb = g.at(blockIdx.y);
    double* rb = r + blockIdx.y * rs;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        rb[i] = (double)gb[i] - rb[i];
}

__global__ void accumulate_kernel(Inst<double> c, const double* d, long ds, long n) {
    double* cb = c.at(blockIdx.y);
    const double* db = d + blockIdx.y * ds;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        cb[i] += db[i];
}

unsigned grid_for(long n) { return (unsigned)std::min<long>((n + 255) / 256, 4096); }

template <class T>
T* dev_alloc(size_t count) {
    return dmalloc_n<T>(count, "table allocation");
}

int4* upload_work(const std::vector<int4>& w, hipStream_t s) {
    int4* d = dev_alloc<int4>(w.size());
    try {
        if (!w.empty())
            check_hip(hipMemcpyAsync(d, w.data(), w.size() * sizeof(int4), hipMemcpyHostToDevice, s),
                      "upload work list");
    } catch (...) {
        dfree(d);
        throw;
    }
    return d;
}

}  // namespace

template <class G>
void grid_residual(Inst<const G> g, double* r, long rs, long n, int nb, hipStream_t s) {
    if (n <= 0 || nb <= 0) return;
    residual_kernel<G><<<dim3(grid_for(n), nb), 256, 0, s>>>(g, r, rs, n);
    check_hip(hipGetLastError(), "residual launch");
}
template void grid_residual<double>(Inst<const double>, double*, long, long, int, hipStream_t);
template void grid_residual<float>(Inst<const float>, double*, long, long, int, hipStream_t);

void coeff_accumulate(Inst<double> c, const double* d, long ds, long n, int nb, hipStream_t s) {
    if (n <= 0 || nb <= 0) return;
    accumulate_kernel<<<dim3(grid_for(n), nb), 256, 0, s>>>(c, d, ds, n);
    check_hip(hipGetLastError(), "accumulate launch");
}

namespace {
// one block per (n, instance): the coefficients of radial degree n are the
// contiguous range [tet(n,n%2,0), tet(n+1,(n+1)%2,0)) of pairs
__global__ void kappa_kernel(Inst<const double> c, double* __restrict__ out, long os, int unnormed) {
    const int n = blockIdx.x;
    const long lo = tet_idx(n, n & 1, 0), hi = tet_idx(n + 1, (n + 1) & 1, 0);
    const double k = unnormed ? (double)(2 * n + 3) : 1.0;
    const double* src = c.at(blockIdx.y);
    double* dst = out + blockIdx.y * os;
    for (long e = 2 * lo + threadIdx.x; e < 2 * hi; e += blockDim.x) dst[e] = k * src[e];
}
}  // namespace

void coeff_kappa_scale(Inst<const double> c, double* out, long os, int nt, int unnormed, int nb,
                       hipStream_t s) {
    if (nt <= 0 || nb <= 0) return;
    kappa_kernel<<<dim3(nt, nb), 256, 0, s>>>(c, out, os, unnormed);
    check_hip(hipGetLastError(), "kappa scale launch");
}

namespace {
// P_lm(z_j) for one (j, m), in the layout [m][l-m][nlat] of the evaluator.
__global__ void eval_plm_kernel(int order, int nlat, int qm, double phase, const double* z,
                                double* ptab) {
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int j = (int)(gid % nlat);
    const int m = (int)(gid / nlat);
    if (m >= order) return;
    const long off = ((long)m * order - ((long)m * (m - 1)) / 2) * nlat;
    plm_column(order, m, z[j], qm, phase, [&](int l, double v) {
        ptab[off + (long)(l - m) * nlat + j] = v;
    });
}
}  // namespace

void radial_table(int order, int nr, int normed, const double* r, double* rtab, hipStream_t s) {
    if (order == 0 || nr == 0) return;
    radial_kernel<<<dim3(nr / 64 + 1, order), 64, 0, s>>>(order, nr, normed, r, rtab);
    check_hip(hipGetLastError(), "radial table");
}

void eval_plm_table(int order, int nlat, int qm, double phase, const double* z, double* ptab,
                    hipStream_t s) {
    if (order == 0 || nlat == 0) return;
    const long threads = (long)order * nlat;
    eval_plm_kernel<<<(unsigned)((threads + 255) / 256), 256, 0, s>>>(order, nlat, qm, phase, z, ptab);
    check_hip(hipGetLastError(), "P table");
}

static long ptab_size(const Geom& g) {
    const int order = g.order;
    return plm_offset(order, order - 1, 1, g.jh) + (long)plm_count(order, order - 1, 1) * g.jh;
}

void ensure_fp32_tables(OrderTables& t, hipStream_t s) {
    if (t.g.order == 0 || t.ptab32 || t.rtab32) return;
    if (t.ptab) {
        const long np = ptab_size(t.g);
        t.ptab32 = dev_alloc<float>(np);
        to_fp32_kernel<<<grid_for(np), 256, 0, s>>>(t.ptab, t.ptab32, np);
    }
    if (t.rtab) {
        const long nr = (long)((t.g.order + 1) * (t.g.order + 1) / 4) * t.g.ncol;
        t.rtab32 = dev_alloc<float>(nr);
        to_fp32_kernel<<<grid_for(nr), 256, 0, s>>>(t.rtab, t.rtab32, nr);
    }
    check_hip(hipGetLastError(), "fp32 tables");
    check_hip(hipStreamSynchronize(s), "fp32 tables");
    if (t.rtab) zt_fused_tables(t, s);
}

void build_tables(OrderTables& t, zgl_kind kind, int zernike_norm, int sh_norm, int sh_phase,
                  int layout, int order, bool want_fp32, hipStream_t s) {
    Geom& g = t.g;
    g.order = order;
    const int lon_min = 2 * order - (order > 0 ? 1 : 0);
    if (kind == ZGL_ZERNIKE) {
        g.nphi = lon_min;
        g.nth = order + 1;
        g.ncol = (order + 1 + 7) / 8 * 8;
        g.lat_stride = g.ncol;
        g.lon_stride = g.nth * g.ncol;
    } else if (layout == ZGL_LAYOUT_LATLON) {
        g.nphi = (lon_min + 7) / 8 * 8;
        g.nth = order;
        g.ncol = 1;
        g.lat_stride = g.nphi;
        g.lon_stride = 1;
    } else {
        g.nphi = lon_min;
        g.nth = (order + 7) / 8 * 8;
        g.ncol = 1;
        g.lat_stride = 1;
        g.lon_stride = g.nth;
    }
    g.jh = (g.nth + 1) / 2;
    g.nm = order;
    if (order == 0) return;

    // Latitude nodes and weights (all of them), then the north half. Each
    // allocation lands in t (freed by free_tables) or in a DeviceTemp at once,
    // so an allocation that fails part way leaks nothing.
    DeviceTemp zt, wt;
    double* z = zt.alloc<double>(g.nth, "table allocation");
    double* w = wt.alloc<double>(g.nth, "table allocation");
    gl_kernel<<<(g.nth + 1) / 2 / 64 + 1, 64, 0, s>>>(g.nth, z, w);
    t.zh = dev_alloc<double>(g.jh);
    double* zh = t.zh;
    t.wlat = dev_alloc<double>(g.jh);
    half_nodes_kernel<<<g.jh / 64 + 1, 64, 0, s>>>(g.nth, z, w, zh, t.wlat);

    // P table, unless the Legendre kernels recompute their P blocks: at high
    // order (the table is O(N^3): 67.6 MB at N = 256 in FP64) when the
    // streaming kernels run; ZGL_PLM=kernel / table forces either.
    t.plm_qm = sh_norm == ZGL_SHNORM_QM ? 1 : 0;
    t.plm_phase = (double)sh_phase;
    {
        const char* e = std::getenv("ZGL_PLM");
        const bool can = kind == ZGL_ZERNIKE && legendre_stream_enabled() &&
                         stream_legendre_fits(order, g.jh);
        const bool want = e && std::strcmp(e, "kernel") == 0 ? true
                        : e && std::strcmp(e, "table") == 0  ? false
                                                             : order >= 192;
        t.plm_in_kernel = can && want;
    }
    if (!t.plm_in_kernel) {
        t.ptab = dev_alloc<double>(ptab_size(g));
        const long threads = (long)order * g.jh;
        plm_kernel<<<(unsigned)((threads + 255) / 256), 256, 0, s>>>(order, g.jh, t.plm_qm,
                                                                    t.plm_phase, zh, t.ptab);
    }

    if (kind == ZGL_ZERNIKE) {
        const int nr = g.ncol;
        DeviceTemp xt, vt;
        double* x = xt.alloc<double>(nr, "table allocation");
        double* v = vt.alloc<double>(nr, "table allocation");
        gl_kernel<<<(nr + 1) / 2 / 64 + 1, 64, 0, s>>>(nr, x, v);
        t.wrad = dev_alloc<double>(nr);
        radial_nodes_kernel<<<nr / 64 + 1, 64, 0, s>>>(nr, x, v, t.wrad);
        const long rsize = (long)((order + 1) * (order + 1) / 4) * nr;
        t.rtab = dev_alloc<double>(rsize);
        radial_kernel<<<dim3(nr / 64 + 1, order), 64, 0, s>>>(
            order, nr, zernike_norm == ZGL_ZNORM_NORMED ? 1 : 0, x, t.rtab);
        check_hip(hipGetLastError(), "table kernels");
        check_hip(hipStreamSynchronize(s), "table generation");
    }
    check_hip(hipGetLastError(), "table kernels");
    check_hip(hipStreamSynchronize(s), "table generation");

    fft_plan_create(t.fft, g.nphi, s);

    // Work lists: (problem, first row of its tile); tiles have 16 * NS rows.
    std::vector<int4> lf, lb, rf, rb;
    const int tm_lf = 16 * kNsLegendreFwd;
    const int tm_lb = 16 * (kind == ZGL_ZERNIKE ? kNsLegendreBwdZt : kNsLegendreBwdSt);
    const int tm_r = 16 * kNsRadial;
    for (int m = 0; m < order; ++m)
        for (int par = 0; par < 2; ++par) {
            const int prob = 2 * m + par;
            const int cnt = plm_count(order, m, par);
            for (int m0 = 0; m0 < cnt; m0 += tm_lf) lf.push_back(make_int4(prob, m0, 0, 0));
            for (int m0 = 0; m0 < g.jh; m0 += tm_lb) lb.push_back(make_int4(prob, m0, 0, 0));
        }
    std::vector<int4> sb, rs;
    if (kind == ZGL_ZERNIKE)
        for (int prob = 0; prob < 2 * order; ++prob)
            for (int m0 = 0; m0 < g.nth / 2; m0 += 16 * kNsStreamBwd) sb.push_back(make_int4(prob, m0, 0, 0));
    if (kind == ZGL_ZERNIKE) {
        for (int l = 0; l < order; ++l) {
            const int ncount = (order - l + 1) / 2;
            for (int m0 = 0; m0 < ncount; m0 += tm_r) rf.push_back(make_int4(l, m0, 0, 0));
            for (int m0 = 0; m0 < 2 * (l + 1); m0 += tm_r) rb.push_back(make_int4(l, m0, 0, 0));
        }
        // radial forward items, costliest first, so that the short items
        // fill in behind the long ones. An item streams its flm rows m <= l
        // (blocks of 16 m) once, whatever its row count, and runs one MFMA
        // strip per 16 rows for each block: cost (blocks of m) x (2 + strips),
        // the flm stream weighted as two strips.
        auto cost = [&](const int4& w) {
            const int ncount = (order - w.x + 1) / 2;
            const int strips = (std::min(tm_r, ncount - w.y) + 15) / 16;
            return (long)((w.x + 16) / 16) * (2 + strips);
        };
        std::stable_sort(rf.begin(), rf.end(), [&](const int4& a, const int4& b) { return cost(a) > cost(b); });
        // radial backward (streaming kernel) items, one per l, costliest
        // first: an item writes its flm rows m <= l (blocks of 16 m) and runs
        // ceil(K / 16) groups of four MFMA steps per block, K = (order - l + 1) / 2
        for (int l = 0; l < order; ++l) rs.push_back(make_int4(l, 0, 0, 0));
        auto cost_b = [&](const int4& w) {
            const int steps = ((order - w.x + 1) / 2 + 15) / 16;
            return (long)((w.x + 16) / 16) * (2 + steps);
        };
        std::stable_sort(rs.begin(), rs.end(), [&](const int4& a, const int4& b) { return cost_b(a) > cost_b(b); });
    }
    t.work_lf = upload_work(lf, s); t.n_lf = (int)lf.size();
    t.work_lb = upload_work(lb, s); t.n_lb = (int)lb.size();
    t.work_rf = upload_work(rf, s); t.n_rf = (int)rf.size();
    t.work_rb = upload_work(rb, s); t.n_rb = (int)rb.size();
    t.work_sb = upload_work(sb, s); t.n_sb = (int)sb.size();
    t.work_rs = upload_work(rs, s); t.n_rs = (int)rs.size();
    check_hip(hipStreamSynchronize(s), "work lists");
    if (want_fp32) ensure_fp32_tables(t, s);
}

void free_tables(OrderTables& t) {
    for (void* p : {(void*)t.wlat, (void*)t.wrad, (void*)t.ptab, (void*)t.zh, (void*)t.rtab, (void*)t.ptab32,
                    (void*)t.rtab32, (void*)t.work_lf, (void*)t.work_lb, (void*)t.work_rf,
                    (void*)t.work_rb, (void*)t.work_sb, (void*)t.work_rs, (void*)t.rfuse_f, (void*)t.rfuse_b,
                    (void*)t.work_ff, (void*)t.work_fb})
        dfree(p);
    fft_plan_destroy(t.fft);
    t = OrderTables{};
}

}  // namespace zgl
