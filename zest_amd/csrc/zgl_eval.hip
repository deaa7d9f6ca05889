// Evaluation of Zernike and spherical-harmonic expansions on arbitrary
// tensor-product grids: zt::GridEvaluator::evaluate and
// st::GridEvaluator::evaluate (include/grid_evaluator.hpp:104-133, 227-265;
// src/grid_evaluator.cpp:44-256).
//
// The same three contractions as the backward transform, at user points:
//   radial     flm[l,m][k] = sum_n R_nl(r_k) c_nlm               (Zernike only)
//   Legendre   S[m][j][k]  = sum_l P_lm(cos theta_j) flm[l,m][k]
//   longitude  v[i][j][k]  = sum_m Re S[m][j][k] cos(m phi_i) + Im S[m][j][k] sin(m phi_i)
// The longitudes are arbitrary, so the last stage is a direct sum (a GEMM
// against the table cos/sin(m phi_i)) instead of an FFT. The radial stage is
// the backward radial kernel run on a table at the given radii; the Legendre
// stage is the streaming kernel with one problem per m (zgl_stream.hip).
//
// Spherical harmonics have one column per point, so a batch is evaluated
// with the instances as the columns: the coefficients are gathered into
// flm[l,m][b] and every stage treats the batch as one wide instance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "zgl_internal.hpp"

namespace zgl {

namespace {

typedef double d2 __attribute__((ext_vector_type(2)));
typedef double d4 __attribute__((ext_vector_type(4)));

// z_j = cos(theta_j); radii padded with zeros to ncol
__global__ void eval_points_kernel(const double* colat, int nlat, const double* r, int nrad,
                                   int ncol, double* z, double* rpad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nlat) z[t] = cos(colat[t]);
    if (rpad && t < ncol) rpad[t] = t < nrad ? r[t] : 0.0;
}

// trig[i][m] = (cos(m phi_i), sin(m phi_i)). m * phi_i is formed exactly as
// p + e (e = fma(m, phi, -p)) and the first-order correction in e applied, so
// the table holds the functions of the given longitudes to about one ulp.
__global__ void eval_trig_kernel(const double* lon, int nlon, int nm, double* trig) {
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (gid >= (long)nlon * nm) return;
    const int i = (int)(gid / nm), m = (int)(gid - (long)i * nm);
    const double phi = lon[i], dm = (double)m;
    const double p = dm * phi;
    const double e = fma(dm, phi, -p);
    double sp, cp;
    sincos(p, &sp, &cp);
    trig[2 * gid] = cp - e * sp;
    trig[2 * gid + 1] = sp + e * cp;
}

// SH: flm[t][b] = coefficient pair t of instance b, zero for b >= n
__global__ void eval_gather_kernel(Inst<const double> c, int n, long ntri, int ncol, double* flm) {
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (gid >= ntri * ncol) return;
    const long t = gid / ncol;
    const int b = (int)(gid - t * ncol);
    d2 v = {0.0, 0.0};
    if (b < n) {
        const double* p = c.at(b) + 2 * t;
        v = d2{p[0], p[1]};
    }
    *reinterpret_cast<d2*>(flm + 2 * gid) = v;
}

// ------------------------------------------------------------ longitude sum
// v[i][col] = sum_m trig[i][m].cos * Re S[m][col] + trig[i][m].sin * Im S[m][col]
// per instance, FP64 MFMA 16x16x4. A workgroup owns 16*kLonMB longitudes and
// 4 * 16*kLonNB columns (each wave 16*kLonNB of them) and walks m in chunks
// of kLonKC: the (cos, sin) block of the chunk is staged in LDS, the S values
// stream from HBM as 16-byte (re, im) loads straight into the B operand, and
// every m step issues two MFMAs per block (cos x re, sin x im) into the same
// accumulator.
constexpr int kLonMB = 8;                 // 128 longitudes per workgroup
constexpr int kLonNB = 2;                 // 32 columns per wave, 128 per workgroup
constexpr int kLonKC = 8;                 // m per LDS chunk
constexpr int kLonRS = 2 * kLonKC + 2;    // LDS row stride (doubles): 16 rows hit distinct banks
constexpr int kLonRows = 16 * kLonMB;
constexpr int kLonCols = 4 * 16 * kLonNB;
constexpr size_t kLonLds = (size_t)kLonRows * kLonRS * sizeof(double);

struct LonArgs {
    const double* trig;           // [nlon][nm][2]
    const double* spec;           // instance base + m * row_stride + 2 * col
    long spec_stride, row_stride; // doubles
    int nm, nlon, ncols;          // complex columns per m row
    int ncol, kvalid;             // col -> (j, k) = (col / ncol, col % ncol); stored if k < kvalid
    double* out;
    long out_inst, out_lon, out_lat, out_k;  // strides (doubles)
    int ntc, ntr;                 // column and row tiles per instance
    long total;                   // tiles (instances x ntc x ntr)
};

__global__ __launch_bounds__(256, 2) void eval_lon_kernel(LonArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* As = reinterpret_cast<double*>(smem_raw);
    // Workgroup b runs on XCD b % 8: each XCD takes a contiguous range of
    // tiles, row tiles fastest, so the workgroups that read one block of S
    // share that XCD's L2.
    const long per = (a.total + 7) / 8;
    const long t = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (t >= a.total) return;
    const int rt = (int)(t % a.ntr);
    const long t2 = t / a.ntr;
    const int ct = (int)(t2 % a.ntc);
    const long inst = t2 / a.ntc;
    const int i0 = rt * kLonRows;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int j = lane & 15, g = lane >> 4;
    const int cw = ct * kLonCols + wave * 16 * kLonNB;
    const double* sb = a.spec + inst * a.spec_stride;

    d4 acc[kLonMB][kLonNB];
#pragma unroll
    for (int q = 0; q < kLonMB; ++q)
#pragma unroll
        for (int c = 0; c < kLonNB; ++c) acc[q][c] = d4{0, 0, 0, 0};

    int colv[kLonNB];
    bool colok[kLonNB];
#pragma unroll
    for (int c = 0; c < kLonNB; ++c) {
        colv[c] = cw + 16 * c + j;
        colok[c] = colv[c] < a.ncols;
    }

    // B values of m step 4s + g: loaded one step ahead, across chunk borders
    auto load = [&](int mb, d2 (&b)[kLonNB]) {
        const int m = mb + g;
#pragma unroll
        for (int c = 0; c < kLonNB; ++c) {
            b[c] = d2{0.0, 0.0};
            if (m < a.nm && colok[c])
                b[c] = *reinterpret_cast<const d2*>(sb + (long)m * a.row_stride + 2L * colv[c]);
        }
    };
    // the (cos, sin) block of a chunk: loaded into registers one chunk ahead,
    // parked in LDS between two barriers
    constexpr int kStage = kLonRows * kLonKC / 256;
    d2 an[kStage];
    auto fetch = [&](int kc) {
#pragma unroll
        for (int u = 0; u < kStage; ++u) {
            const int e = tid + 256 * u;
            const int r = e / kLonKC, mm = e - r * kLonKC;
            const int i = i0 + r, m = kc + mm;
            an[u] = d2{0.0, 0.0};
            if (i < a.nlon && m < a.nm) an[u] = *reinterpret_cast<const d2*>(a.trig + ((long)i * a.nm + m) * 2);
        }
    };
    d2 bn[kLonNB];
    load(0, bn);
    fetch(0);
    for (int kc = 0; kc < a.nm; kc += kLonKC) {
        __syncthreads();  // the previous chunk's A reads are done
#pragma unroll
        for (int u = 0; u < kStage; ++u) {
            const int e = tid + 256 * u;
            const int r = e / kLonKC, mm = e - r * kLonKC;
            *reinterpret_cast<d2*>(As + r * kLonRS + 2 * mm) = an[u];
        }
        __syncthreads();
        if (kc + kLonKC < a.nm) fetch(kc + kLonKC);
        const int steps = (min(kLonKC, a.nm - kc) + 3) >> 2;
#pragma unroll
        for (int s = 0; s < kLonKC / 4; ++s) {
            if (s < steps) {
                d2 b[kLonNB];
#pragma unroll
                for (int c = 0; c < kLonNB; ++c) b[c] = bn[c];
                const int next = kc + 4 * (s + 1);
                if (next < a.nm) load(next, bn);
                d2 x[kLonMB];
#pragma unroll
                for (int q = 0; q < kLonMB; ++q)
                    x[q] = *reinterpret_cast<const d2*>(As + (16 * q + j) * kLonRS + 2 * (4 * s + g));
                // all cos products, then all sin products: no MFMA waits on
                // the one just issued
#pragma unroll
                for (int q = 0; q < kLonMB; ++q)
#pragma unroll
                    for (int c = 0; c < kLonNB; ++c)
                        acc[q][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[q][0], b[c][0], acc[q][c], 0, 0, 0);
#pragma unroll
                for (int q = 0; q < kLonMB; ++q)
#pragma unroll
                    for (int c = 0; c < kLonNB; ++c)
                        acc[q][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[q][1], b[c][1], acc[q][c], 0, 0, 0);
            }
        }
    }

    double* ob = a.out + inst * a.out_inst;
#pragma unroll
    for (int c = 0; c < kLonNB; ++c) {
        if (!colok[c]) continue;
        const int jj = colv[c] / a.ncol, k = colv[c] - jj * a.ncol;
        if (k >= a.kvalid) continue;
        double* oc = ob + jj * a.out_lat + k * a.out_k;
#pragma unroll
        for (int q = 0; q < kLonMB; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + 16 * q + g + 4 * r;  // D row of this lane
                if (i < a.nlon) oc[i * a.out_lon] = acc[q][c][r];
            }
    }
}

void lon_sum(LonArgs a, int ninst, hipStream_t s) {
    a.ntc = (a.ncols + kLonCols - 1) / kLonCols;
    a.ntr = (a.nlon + kLonRows - 1) / kLonRows;
    a.total = (long)ninst * a.ntc * a.ntr;
    if (a.total == 0) return;
    const long blocks = (a.total + 7) / 8 * 8;
    check_hip(hipFuncSetAttribute(reinterpret_cast<const void*>(&eval_lon_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLonLds),
              "LDS attribute");
    eval_lon_kernel<<<(unsigned)blocks, 256, kLonLds, s>>>(a);
    check_hip(hipGetLastError(), "longitude sum launch");
}

template <class T>
void grow(T*& p, size_t& cap, size_t need) {
    if (need <= cap) return;
    if (p) check_hip(hipDeviceSynchronize(), "scratch growth");
    dfree(p);
    p = nullptr;
    cap = 0;
    p = dmalloc_n<T>(need, "scratch allocation");
    cap = need;
}

int4* upload(const std::vector<int4>& w, hipStream_t s) {
    int4* d = dmalloc_n<int4>(w.size(), "work list");
    try {
        if (!w.empty()) {
            check_hip(hipMemcpyAsync(d, w.data(), w.size() * sizeof(int4), hipMemcpyHostToDevice, s),
                      "work list");
            check_hip(hipStreamSynchronize(s), "work list");
        }
    } catch (...) {
        dfree(d);
        throw;
    }
    return d;
}

// P recomputed in the Legendre kernel only when its table would be large:
// the evaluator's table is read once per call, while each workgroup would
// otherwise rerun the column recursion for its points.
bool plm_in_kernel(long table_doubles) {
    const char* e = std::getenv("ZGL_PLM");
    if (e && std::strcmp(e, "kernel") == 0) return true;
    if (e && std::strcmp(e, "table") == 0) return false;
    return table_doubles * sizeof(double) > (size_t(1) << 30);
}

constexpr size_t kEvalScratchBudget = size_t(2) << 30;  // bytes of flm + S per chunk
constexpr int kShColumns = 64;                          // SH instances per chunk

}  // namespace

void EvalCache::release() {
    for (double* p : {pts, ptab, rtab, trig, flm, spec}) dfree(p);
    for (auto& kv : leg_work) dfree(kv.second.first);
    for (auto& kv : rad_work) dfree(kv.second.first);
    *this = EvalCache{};
}

void evaluate(EvalCache& ec, const EvalCall& e, hipStream_t s, const EvalTimer& timer) {
    auto run = [&](int stage, int nb, const std::function<void()>& f) {
        if (timer) timer(stage, nb, f);
        else f();
    };
    const bool zt = e.kind == ZGL_ZERNIKE;
    const int N = e.order;
    const long npts = (long)e.nlon * e.nlat * (zt ? e.nrad : 1);
    if (npts == 0 || e.nb == 0) return;
    if (N == 0) {
        check_hip(hipMemset2DAsync(e.values, e.value_stride * sizeof(double), 0, npts * sizeof(double),
                                   e.nb, s),
                  "zero values");
        return;
    }
    const int ns = eval_legendre_strips(N);
    const long ntri = (long)N * (N + 1) / 2;

    // columns per point (Zernike: padded radii) and instances per chunk
    int ncol, cb;
    if (zt) {
        ncol = (e.nrad + 7) / 8 * 8;
        const double per = 16.0 * ((double)ntri * ncol + (double)N * e.nlat * ncol);
        cb = (int)std::max(1.0, std::min((double)e.nb, (double)kEvalScratchBudget / per));
    } else {
        cb = std::min(e.nb, kShColumns);
        ncol = (cb + 7) / 8 * 8;
    }
    // the Legendre kernel addresses one instance's flm and S with 32-bit offsets
    if (2.0 * ntri * ncol >= 4294967296.0 || 2.0 * N * e.nlat * ncol >= 4294967296.0 ||
        (double)e.nlat * ncol >= 2147483648.0)
        throw Error{ZGL_ERROR_NOT_SUPPORTED, "evaluation grid too large for one call: split the points"};
    const long flm_stride = 2 * ntri * ncol;            // doubles per instance
    const long spec_stride = 2L * N * e.nlat * ncol;
    const int ninst = zt ? cb : 1;                      // instances of the staged arrays

    // point tables
    grow(ec.pts, ec.pts_cap, (size_t)e.nlat + ncol);
    double* z = ec.pts;
    double* rpad = ec.pts + e.nlat;
    const bool in_kernel = plm_in_kernel(ntri * e.nlat);
    if (!in_kernel) grow(ec.ptab, ec.ptab_cap, (size_t)ntri * e.nlat);
    grow(ec.trig, ec.trig_cap, (size_t)2 * e.nlon * N);
    if (zt) grow(ec.rtab, ec.rtab_cap, (size_t)((N + 1) * (N + 1) / 4) * ncol);
    run(0, e.nb, [&] {
        const int tp = std::max(e.nlat, ncol);
        eval_points_kernel<<<(tp + 255) / 256, 256, 0, s>>>(e.colat, e.nlat, e.radii, e.nrad, ncol, z,
                                                           zt ? rpad : nullptr);
        check_hip(hipGetLastError(), "points");
        if (!in_kernel) eval_plm_table(N, e.nlat, e.qm, e.phase, z, ec.ptab, s);
        const long n = (long)e.nlon * N;
        eval_trig_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(e.lon, e.nlon, N, ec.trig);
        check_hip(hipGetLastError(), "trig table");
        if (zt) radial_table(N, ncol, e.znorm == ZGL_ZNORM_NORMED ? 1 : 0, rpad, ec.rtab, s);
    });
    OrderTables et{};
    if (zt) {
        auto it = ec.rad_work.find(N);
        if (it == ec.rad_work.end()) {
            std::vector<int4> rb;
            for (int l = 0; l < N; ++l)
                for (int m0 = 0; m0 < 2 * (l + 1); m0 += 16 * kNsRadial) rb.push_back(make_int4(l, m0, 0, 0));
            it = ec.rad_work.emplace(N, std::make_pair(upload(rb, s), (int)rb.size())).first;
        }
        et.g.order = N;
        et.g.ncol = ncol;
        et.rtab = ec.rtab;
        et.work_rb = it->second.first;
        et.n_rb = it->second.second;
    }
    const auto key = std::make_tuple(N, e.nlat, ns);
    auto lw = ec.leg_work.find(key);
    if (lw == ec.leg_work.end() && ec.leg_work.size() >= 64) {
        // many distinct point counts: start the cache over
        check_hip(hipDeviceSynchronize(), "work list cache");
        for (auto& kv : ec.leg_work) dfree(kv.second.first);
        ec.leg_work.clear();
        lw = ec.leg_work.end();
    }
    if (lw == ec.leg_work.end()) {
        std::vector<int4> w;
        for (int m = 0; m < N; ++m)
            for (int m0 = 0; m0 < e.nlat; m0 += 16 * ns) w.push_back(make_int4(m, m0, 0, 0));
        lw = ec.leg_work.emplace(key, std::make_pair(upload(w, s), (int)w.size())).first;
    }

    grow(ec.flm, ec.flm_cap, (size_t)flm_stride * ninst);
    grow(ec.spec, ec.spec_cap, (size_t)spec_stride * ninst);

    for (int b0 = 0; b0 < e.nb; b0 += cb) {
        const int n = std::min(cb, e.nb - b0);
        const Inst<const double> ci = e.coeffs.from(b0);
        run(1, n, [&] {
            if (zt) {
                ZtGemmArgs ra{&et, N, 0, nullptr, 0, ec.flm, flm_stride / 2,
                              Inst<double>{const_cast<double*>(ci.base), ci.stride, ci.offs}, n, 0};
                zt_radial_backward<double>(ra, s);
            } else {
                const long cnt = ntri * ncol;
                eval_gather_kernel<<<(unsigned)((cnt + 255) / 256), 256, 0, s>>>(ci, n, ntri, ncol, ec.flm);
                check_hip(hipGetLastError(), "gather");
            }
        });
        EvalLegArgs la{in_kernel ? nullptr : ec.ptab, z, e.qm, e.phase, ec.flm, flm_stride,
                       ec.spec, spec_stride, N, N, e.nlat, ncol, zt ? n : 1,
                       lw->second.first, lw->second.second, ns};
        run(2, n, [&] {
            if (!eval_legendre(la, s)) throw Error{ZGL_ERROR_NOT_SUPPORTED, "evaluator Legendre stage"};
        });
        LonArgs lo{};
        lo.trig = ec.trig;
        lo.spec = ec.spec;
        lo.spec_stride = spec_stride;
        lo.row_stride = 2L * e.nlat * ncol;
        lo.nm = N;
        lo.nlon = e.nlon;
        lo.ncols = e.nlat * ncol;
        lo.ncol = ncol;
        if (zt) {
            lo.kvalid = e.nrad;
            lo.out = e.values + (long)b0 * e.value_stride;
            lo.out_inst = e.value_stride;
            lo.out_lon = (long)e.nlat * e.nrad;
            lo.out_lat = e.nrad;
            lo.out_k = 1;
        } else {
            lo.kvalid = n;
            lo.out = e.values + (long)b0 * e.value_stride;
            lo.out_inst = 0;
            lo.out_lon = e.nlat;
            lo.out_lat = 1;
            lo.out_k = e.value_stride;
        }
        run(3, n, [&] { lon_sum(lo, zt ? n : 1, s); });
    }
}

}  // namespace zgl
