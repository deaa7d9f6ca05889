// Internal declarations shared by the zgl HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "zgl.h"

namespace zgl {

constexpr int kMaxPass = 16;

// Error carrying a zgl_status across internal helpers.
struct Error {
    zgl_status status;
    std::string what;
};

void check_hip(hipError_t e, const char* what);  // throws Error

// Every device allocation of the library: through the allocator installed
// with zgl_set_device_allocator, else hipMalloc. dmalloc throws Error
// (ZGL_ERROR_OUT_OF_MEMORY when the memory is not there); dfree returns the
// memory to the allocator that made it, after the device has finished the
// work queued on it, and ignores NULL.
void* dmalloc(size_t bytes, const char* what);
void dfree(void* p) noexcept;

template <class T>
T* dmalloc_n(size_t count, const char* what) {
    return static_cast<T*>(dmalloc((count ? count : 1) * sizeof(T), what));
}

// A device buffer that is freed when it goes out of scope (temporaries of
// table generation: an exception on the way leaks nothing).
struct DeviceTemp {
    void* p = nullptr;
    DeviceTemp() = default;
    DeviceTemp(const DeviceTemp&) = delete;
    DeviceTemp& operator=(const DeviceTemp&) = delete;
    ~DeviceTemp() { dfree(p); }
    template <class T>
    T* alloc(size_t count, const char* what) {
        dfree(p);
        p = nullptr;
        p = dmalloc_n<T>(count, what);
        return static_cast<T*>(p);
    }
};

// Mixed-radix DIT plan of one real-DFT length (device tables + pass list).
struct FftDesc {
    int n;
    int npass;
    int radix[kMaxPass];  // radix of each pass, in execution order
    int L[kMaxPass];      // product of the radices before the pass
};

struct FftPlan {
    FftDesc desc{};
    double2* tw = nullptr;           // e^{-2 pi i t / n}, t < n
    float2* tw32 = nullptr;
    double2* twp = nullptr;          // input twiddle of every slot, passes 1..npass-1
    float2* twp32 = nullptr;
    unsigned short* perm = nullptr;  // LDS slot of input sample i
    int tile[2] = {16, 16};          // column pairs per workgroup: FP64, FP32
    bool mfma_pass = false;          // a prime radix runs on the matrix cores
    bool lean = false;               // ... and every other radix is at most 7
    bool wide = false;               // a lean plan with a prime radix in (255, 511]: 512 threads
};

// Geometry of one order for a plan, both transform kinds.
struct Geom {
    int order = 0;
    int nphi = 0;   // FFT length (longitudes)
    int nth = 0;    // latitudes
    int ncol = 1;   // columns per latitude (radial nodes for Zernike, 1 for SH)
    int jh = 0;     // half-latitude count (nth + 1) / 2
    int nm = 0;     // longitudinal modes kept (= order)
    int lat_stride = 0;  // distance between latitudes in the grid (doubles)
    int lon_stride = 0;  // distance between longitudes in the grid (doubles)
};

// Device-resident tables of one order.
struct OrderTables {
    Geom g;
    double* wlat = nullptr;  // [jh] weights of the north-half latitude nodes
    double* wrad = nullptr;  // [ncol] r^2 v of the radial nodes (Zernike)
    double* ptab = nullptr;  // P_lm(z_h), parity-split [m][par][l'][h]; null when the
                             // Legendre kernels recompute their P blocks (plm_in_kernel)
    double* zh = nullptr;    // [jh] north-half latitude nodes z_h
    int plm_qm = 0;          // P_lm conventions: qm normalisation, phase (+1 / -1)
    double plm_phase = -1.0;
    bool plm_in_kernel = false;
    double* rtab = nullptr;  // R_nl(r_k), zest layout [((n+1)^2>>2)+(l>>1)][nr]
    float* ptab32 = nullptr; // FP32 roundings of the two tables (FP32 modes)
    float* rtab32 = nullptr;
    FftPlan fft;
    int4* work_lf = nullptr; int n_lf = 0;  // (problem, m0) tiles, forward Legendre
    int4* work_lb = nullptr; int n_lb = 0;  // backward Legendre
    int4* work_rf = nullptr; int n_rf = 0;  // forward radial
    int4* work_rb = nullptr; int n_rb = 0;  // backward radial
    int4* work_sb = nullptr; int n_sb = 0;  // backward Legendre, streaming kernel (Zernike)
    int4* work_rs = nullptr; int n_rs = 0;  // backward radial, streaming kernel: l, costliest first
    // fused FP32 colatitude + radial kernels (Zernike, orders <= 128): R_nl(r_k)
    // by 16-node chunk, [l][kc][d][16] (forward) and [l][kc][16][d] (backward),
    // n = l + 2d, d < fuse_dp; (m, par) problems, costliest first
    float* rfuse_f = nullptr;
    float* rfuse_b = nullptr;
    int fuse_nkc = 0, fuse_dp = 0;
    int2* work_ff = nullptr; int n_ff = 0;
    int2* work_fb = nullptr; int n_fb = 0;
};

// Rows per GEMM tile are 16 * NS; the work lists step by that many rows.
constexpr int kNsLegendreFwd = 4;
constexpr int kNsLegendreBwdZt = 5;  // 80 rows: one tile for 65 half latitudes
constexpr int kNsLegendreBwdSt = 4;
constexpr int kNsRadial = 4;
// Streaming backward Legendre: 16 * kNsStreamBwd rows of the nth/2 latitudes
// off the equator per tile; the equator row (odd nth) is computed beside them.
constexpr int kNsStreamBwd = 4;

// Offsets of the parity-split P table: rows l = m + par + 2 l' < order.
__host__ __device__ inline int plm_count(int order, int m, int par) {
    const int r = order - m - par;
    return r > 0 ? (r + 1) / 2 : 0;
}
__host__ __device__ inline long plm_offset(int order, int m, int par, int jh) {
    long rows = (long)m * order - ((long)m * (m - 1)) / 2;  // sum_{m'<m} (order - m')
    if (par) rows += plm_count(order, m, 0);
    return rows * jh;
}
__host__ __device__ inline long tri_idx(long l, long m) { return ((l * (l + 1)) >> 1) + m; }
__host__ __device__ inline long rad_idx(long n, long l) { return (((n + 1) * (n + 1)) >> 2) + (l >> 1); }
__host__ __device__ inline long tet_idx(long n, long l, long m) {
    return (n + 1) * (n + 3) * (2 * n + 1) / 24 + ((l * l) >> 2) + m;
}

// ---- table generation (zgl_tables.hip) ----
void build_tables(OrderTables& t, zgl_kind kind, int zernike_norm, int sh_norm,
                  int sh_phase, int layout, int order, bool want_fp32, hipStream_t s);
void ensure_fp32_tables(OrderTables& t, hipStream_t s);
void free_tables(OrderTables& t);

// ---- FFT (zgl_fft.hip) ----
void fft_plan_create(FftPlan& p, int n, hipStream_t s);
void fft_plan_destroy(FftPlan& p);

// Addressing of the instances of a batch: base + b * stride, or, for a
// ragged batch (instances anywhere in memory), base + offs[b] with offs a
// device array of element offsets from base.
template <class P>
struct Inst {
    P* base = nullptr;
    long stride = 0;
    const long* offs = nullptr;
    __host__ __device__ P* at(long b) const { return base + (offs ? offs[b] : b * stride); }
    // the instances from b0 on
    Inst from(long b0) const {
        return offs ? Inst{base, stride, offs + b0} : Inst{base + b0 * stride, stride, nullptr};
    }
    operator Inst<const P>() const { return Inst<const P>{base, stride, offs}; }
};

enum GridLayoutKind { GRID_COLS = 0, GRID_ROWS = 1 };
// GRID_COLS: sample i of latitude j, column c at i*lon_stride + j*lat_stride + c
//            (Zernike LonLatRad and SH LonLat; the transform axis is the slow one).
// GRID_ROWS: sample i of latitude j at j*lat_stride + i (SH LatLon).

struct FftLaunch {
    const Geom* g;
    const OrderTables* t;
    int layout_kind;
    int nm_used;            // modes written (forward) or read (backward)
    bool weighted;          // forward: quadrature weights on the input (off for the
                            // transpose of the backward); backward: weights on the
                            // output (on only for the transpose of the forward)
    double scale;           // prefactor applied with the weights
    double center_factor;   // weight factor of an equator node (zest's SH quirk: 2)
    int tile = 0;           // column pairs per workgroup (0: the plan's choice)
    bool grid_aligned16 = false;  // every instance's grid base is 16-byte aligned
};
// Tile widths the FFT kernels are built for that fit the LDS at this length
// and arithmetic (widest first).
std::vector<int> fft_tiles(const FftPlan& p, bool fp32);

// R = double or float: the arithmetic; G: the grid element type (double, or
// float with R = float). spec holds complex R values and its stride counts
// complex elements.
template <class R, class G>
void fft_forward(const FftLaunch& L, Inst<const G> grid, void* spec, long spec_stride, int nb,
                 hipStream_t s);
template <class R, class G>
void fft_backward(const FftLaunch& L, const void* spec, long spec_stride, Inst<G> grid, int nb,
                  hipStream_t s);

// ---- contractions (zgl_gemm.hip) ----
struct ZtGemmArgs {
    const OrderTables* t;
    int nt;                 // truncated order min(exp_order, grid_order)
    int zernike_unnormed;   // scale coefficients by 2n+3
    void* spec; long spec_stride;   // complex elements of the compute type
    void* flm; long flm_stride;
    Inst<double> coeffs;    // strides in doubles
    int nb;
    int group = 0;          // instances per workgroup of the streaming kernels (0: default);
                            // Legendre stages: | kGroupX3 for the BF16x3 products (FP32),
                            // or kGroupTiles | tiles per wave (runs of equal tile counts)
};
constexpr int kGroupX3 = 16;
constexpr int kGroupTiles = 32;
template <class T> void zt_legendre_forward(const ZtGemmArgs& a, hipStream_t s);
template <class T> void zt_radial_forward(const ZtGemmArgs& a, hipStream_t s);
template <class T> void zt_radial_backward(const ZtGemmArgs& a, hipStream_t s);
template <class T> void zt_legendre_backward(const ZtGemmArgs& a, hipStream_t s);
// Streaming forms of the Legendre and radial stages (zgl_stream.hip): false
// when the table block of a workgroup does not fit the LDS budget (large
// orders), in which case nothing was launched.
template <class T> bool zt_legendre_forward_stream(const ZtGemmArgs& a, hipStream_t s);
template <class T> bool zt_legendre_backward_stream(const ZtGemmArgs& a, hipStream_t s);
template <class T> bool zt_radial_forward_stream(const ZtGemmArgs& a, hipStream_t s);
template <class T> bool zt_radial_backward_stream(const ZtGemmArgs& a, hipStream_t s);
// Fused FP32 colatitude + radial stages (zgl_fused.hip), orders <= 128:
// false when the order does not fit (nothing launched); on only with
// ZGL_FUSED=1 at table build and call time.
bool zt_fused_fits(const OrderTables& t);
void zt_fused_tables(OrderTables& t, hipStream_t s);
bool zt_fused_forward(const ZtGemmArgs& a, hipStream_t s);
bool zt_fused_backward(const ZtGemmArgs& a, hipStream_t s);
// ZGL_LEGENDRE_ENGINE=tiled selects the tiled LDS engine for every order and
// stage.
bool legendre_stream_enabled();
// Whether both streaming Legendre kernels run at this order in FP64 (their P
// blocks fit the LDS budget), which in-kernel P_lm recomputation requires.
bool stream_legendre_fits(int order, int jh);

struct StGemmArgs {
    const OrderTables* t;
    int nt;
    void* spec; long spec_stride;
    Inst<double> coeffs;    // strides in doubles
    int nb;
    void* stage;            // coefficient-major staging, nt(nt+1)/2 x nb complex
};
template <class T> void st_legendre_forward(const StGemmArgs& a, hipStream_t s);
template <class T> void st_legendre_backward(const StGemmArgs& a, hipStream_t s);

// ---- evaluation at arbitrary points (zgl_eval.hip, zgl_stream.hip) ----
// S[inst][m][j][c] = sum_{l=m}^{nt-1} P_lm(z_j) flm[inst][l,m][c], complex
// columns c < ncol; P from the table [m][l-m][nlat] (rows for l < order) or,
// with ptab null, recomputed from z in the kernel. Work items (m, first
// point) step by 16 * ns points.
struct EvalLegArgs {
    const double* ptab;
    const double* z;
    int qm; double phase;
    const double* flm; long flm_stride;  // doubles per instance
    double* spec; long spec_stride;      // doubles per instance
    int order, nt, nlat, ncol, nb;
    const int4* work; int nwork; int ns;
};
bool eval_legendre(const EvalLegArgs& a, hipStream_t s);
// The 16-point strips per workgroup (4 or 2) whose P block of every l fits
// the LDS budget at this order, else 1: one strip, the l range summed in
// passes that add to S.
int eval_legendre_strips(int nt);
// R_nl(r_k) table [((n+1)^2>>2)+(l>>1)][nr] at nr given radii (device).
void radial_table(int order, int nr, int normed, const double* r, double* rtab, hipStream_t s);
// P_lm(z_j) table [m][l-m][nlat] at nlat given z (device).
void eval_plm_table(int order, int nlat, int qm, double phase, const double* z, double* ptab,
                    hipStream_t s);

// One evaluation call: nb expansions of one order at the tensor-product grid
// lon x colat (x radii for Zernike). All pointers are device memory; values of
// instance b start at values + b * value_stride, shaped [nlon][nlat][nrad]
// (Zernike) or [nlon][nlat] (SH).
struct EvalCall {
    zgl_kind kind;
    int znorm, qm;
    double phase;
    int order;
    Inst<const double> coeffs;
    const double* lon; int nlon;
    const double* colat; int nlat;
    const double* radii; int nrad;
    double* values; long value_stride;
    int nb;
};
// Device buffers and work lists of the evaluator, kept between calls.
struct EvalCache {
    double* pts = nullptr; size_t pts_cap = 0;    // doubles: z, padded radii
    double* ptab = nullptr; size_t ptab_cap = 0;
    double* rtab = nullptr; size_t rtab_cap = 0;
    double* trig = nullptr; size_t trig_cap = 0;
    double* flm = nullptr; size_t flm_cap = 0;
    double* spec = nullptr; size_t spec_cap = 0;
    std::map<std::tuple<int, int, int>, std::pair<int4*, int>> leg_work;  // (order, nlat, strips)
    std::map<int, std::pair<int4*, int>> rad_work;                      // order
    void release();
};
// Optional stage brackets (the plan's stage timing): timer(stage, instances,
// launch) runs `launch`; stages 0 tables, 1 radial, 2 Legendre, 3 longitude.
using EvalTimer = std::function<void(int, int, const std::function<void()>&)>;
void evaluate(EvalCache& ec, const EvalCall& e, hipStream_t s, const EvalTimer& timer = nullptr);

// ---- power spectra (zgl_spectra.hip) ----
// Entries per instance: ((order+1)^2)/4 for Zernike ([rad_idx(n,l)]), order for SH.
long spectrum_size(zgl_kind kind, int order);
// out[b][e] = norm * sum over the m of entry e of a_b . b_b ({cos, sin} pairs,
// the sin part of m = 0 excluded); b = a for the power spectrum.
void power_spectrum(zgl_kind kind, int sh_norm, int order, Inst<const double> a,
                    Inst<const double> b, double* out, long out_stride, int nb, hipStream_t s);

// ---- rotations (zgl_rotate.hip) ----
// Doubles of the Wigner d(pi/2) table of degrees l < order, zest's layout
// [l(l+1)(2l+1)/6 + (l+1) m1 + m2].
long wignerd_size(int order);
void wignerd_table(int order, double* d, hipStream_t s);
// Work items of a batch of nb: (l, column tile) for the fused kernel (l <=
// 255), (l, row parity, row tile, column tile) for the two passes (l > 255).
std::vector<int4> rotation_fused_work(bool zt, int order, int nb);
std::vector<int4> rotation_work(bool zt, int order, int nb);
// In-place ZYZ rotation of the degrees l <= 255 of nb expansions; angles
// [nb][3] (device).
void rotate_fused(bool zt, int order, const double* dtab, const int4* work, int nwork, Inst<double> coeffs,
                  const double* angles, int coordinate, int nb, hipStream_t s);
// One of the two passes of the in-place ZYZ rotation of the degrees l > 255 (pass 0:
// coefficients -> temp through X(+90); pass 1: temp -> coefficients through
// X(-90)); angles [nb][3] (device); temp holds nb complex copies of the
// coefficient layout (temp_stride doubles apart).
void rotate_pass(int pass, bool zt, int order, const double* dtab, const int4* work, int nwork,
                 Inst<double> coeffs, double* temp, long temp_stride, const double* angles, int coordinate,
                 int nb, hipStream_t s);
// In-place rotation about z; angles [nb] (device).
void polar_rotate(bool zt, int order, Inst<double> coeffs, const double* angles, int coordinate, int nb,
                  hipStream_t s);

// ---- elementwise helpers (zgl_tables.hip) ----
// r[b][i] = g[b][i] - r[b][i] for i < n, instance strides in elements.
template <class G>
void grid_residual(Inst<const G> g, double* r, long rs, long n, int nb, hipStream_t s);
// c[b][i] += d[b][i] for i < n.
void coeff_accumulate(Inst<double> c, const double* d, long ds, long n, int nb, hipStream_t s);
// Zernike coefficients times kappa_n = 2n+3 (when `unnormed`), n < nt:
// out[b][tet(n,l,m)] = kappa_n * c[b][tet(n,l,m)] (both parts).
void coeff_kappa_scale(Inst<const double> c, double* out, long os, int nt, int unnormed, int nb,
                       hipStream_t s);

// ---- persistent launches over many orders (ragged batches) ----
// One launch of a stage kernel over groups of instances of different orders:
// the arguments of every group in a device array, and the work items
// (group, x, y) -- x, y being the stage's own workgroup coordinates within the
// group -- dealt to a fixed number of resident workgroups, longest first, each
// walking its own contiguous range of items. `run` enqueues the launch; the
// device arrays stay valid until release().
struct PersistentLaunch {
    std::vector<void*> buffers;
    std::function<void(hipStream_t)> run;
    const char* stage = "";
    long items = 0;
    long instances = 0;  // instances the launch covers (stage statistics)
    int lane = 0;        // internal stream it runs on (ragged pipelines)
    int blocks = 0;
    double cost_max = 0, cost_mean = 0;  // modelled work of the busiest and of the mean workgroup
    std::vector<double> bin_cost;        // modelled work of each workgroup
};
void release(PersistentLaunch& pl);
void release(std::vector<PersistentLaunch>& v);

// Deals the items, with their modelled costs, to nbins bins by longest
// processing time first (each to the least loaded bin); returns them bin by
// bin in `dealt` with the nbins + 1 bin offsets in `bounds`.
void lpt_deal(const std::vector<int4>& items, const std::vector<double>& cost, int nbins,
              std::vector<int4>& dealt, std::vector<int>& bounds, PersistentLaunch& pl);
// Copies v to new device memory owned by pl (synchronous).
void* upload_owned(PersistentLaunch& pl, const void* data, size_t bytes);
template <class T>
T* upload_owned(PersistentLaunch& pl, const std::vector<T>& v) {
    return static_cast<T*>(upload_owned(pl, v.data(), v.size() * sizeof(T)));
}
// Resident workgroups of a kernel on the current device (CUs x per-CU occupancy).
int resident_blocks(const void* kernel, int threads, size_t lds);

// The FFT of one group: FftLaunch as for a batched call, the group's grids
// (Inst) and its spectrum block on the wave's scratch.
struct FftGroup {
    FftLaunch L;
    Inst<double> grid;
    void* spec;
    long spec_stride;
    int nb;
};
// Appends the persistent launches of one FFT direction (one per kernel
// variant among the groups) to out. R: the arithmetic; grids are FP64.
// False (nothing appended) for groups whose FFT width has no persistent form.
template <class R> bool fft_forward_persistent(std::vector<PersistentLaunch>& out, const std::vector<FftGroup>& g);
template <class R> bool fft_backward_persistent(std::vector<PersistentLaunch>& out, const std::vector<FftGroup>& g);
// The streaming Legendre and radial stages (zgl_stream.hip); false when a
// group does not fit the streaming kernels (nothing appended then).
template <class T> bool zt_legendre_forward_persistent(std::vector<PersistentLaunch>& out,
                                                       const std::vector<ZtGemmArgs>& g);
template <class T> bool zt_legendre_backward_persistent(std::vector<PersistentLaunch>& out,
                                                        const std::vector<ZtGemmArgs>& g);
template <class T> bool zt_radial_forward_persistent(std::vector<PersistentLaunch>& out,
                                                     const std::vector<ZtGemmArgs>& g);
template <class T> bool zt_radial_backward_persistent(std::vector<PersistentLaunch>& out,
                                                      const std::vector<ZtGemmArgs>& g);

}  // namespace zgl
