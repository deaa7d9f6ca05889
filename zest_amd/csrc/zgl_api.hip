// The C ABI of include/zgl.h: plans, table caching, chunked pipelines and
// host/device staging around the HIP kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>
#include <new>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "zgl_internal.hpp"

namespace zgl {

void check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    zgl_status st = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
                        ? ZGL_ERROR_OUT_OF_MEMORY
                        : ZGL_ERROR_HIP;
    throw Error{st, std::string(what) + ": " + hipGetErrorString(e)};
}

namespace {

// The allocator installed by zgl_set_device_allocator, and the live
// allocations made through one (with the allocator that made each, so that
// it gets its memory back whatever is installed by then).
struct DeviceAllocator {
    zgl_device_alloc_fn alloc = nullptr;
    zgl_device_free_fn free = nullptr;
    void* user = nullptr;
};
struct HookedBlock {
    DeviceAllocator by;
    size_t bytes;
};
std::mutex g_alloc_mutex;
DeviceAllocator g_allocator;
std::unordered_map<void*, HookedBlock> g_hooked;

}  // namespace

void* dmalloc(size_t bytes, const char* what) {
    DeviceAllocator a;
    {
        std::lock_guard<std::mutex> lock(g_alloc_mutex);
        a = g_allocator;
    }
    if (!a.alloc) {
        void* p = nullptr;
        check_hip(hipMalloc(&p, bytes), what);
        return p;
    }
    void* p = a.alloc(bytes, a.user);
    if (!p)
        throw Error{ZGL_ERROR_OUT_OF_MEMORY,
                    std::string(what) + ": the installed device allocator returned NULL for " +
                        std::to_string(bytes) + " bytes"};
    if (reinterpret_cast<uintptr_t>(p) % 256 != 0) {
        a.free(p, bytes, a.user);
        throw Error{ZGL_ERROR_INVALID_VALUE,
                    std::string(what) + ": the installed device allocator returned memory not aligned to 256 bytes"};
    }
    try {
        std::lock_guard<std::mutex> lock(g_alloc_mutex);
        g_hooked.emplace(p, HookedBlock{a, bytes});
    } catch (...) {
        a.free(p, bytes, a.user);
        throw;
    }
    return p;
}

void release(PersistentLaunch& pl) {
    for (void* b : pl.buffers) dfree(b);
    pl = PersistentLaunch{};
}

void release(std::vector<PersistentLaunch>& v) {
    for (auto& pl : v) release(pl);
    v.clear();
}

void lpt_deal(const std::vector<int4>& items, const std::vector<double>& cost, int nbins,
              std::vector<int4>& dealt, std::vector<int>& bounds, PersistentLaunch& pl) {
    const size_t n = items.size();
    nbins = std::max(1, std::min<int>(nbins, (int)std::max<size_t>(n, 1)));
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    // min-heap of (load, bin)
    std::vector<std::pair<double, int>> heap;
    heap.reserve(nbins);
    for (int b = 0; b < nbins; ++b) heap.push_back({0.0, b});
    auto cmp = [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
        return a.first > b.first || (a.first == b.first && a.second > b.second);
    };
    std::vector<int> bin_of(n);
    std::vector<double> load(nbins, 0.0);
    for (size_t k : order) {
        std::pop_heap(heap.begin(), heap.end(), cmp);
        auto& top = heap.back();
        bin_of[k] = top.second;
        top.first += cost[k];
        load[top.second] = top.first;
        std::push_heap(heap.begin(), heap.end(), cmp);
    }
    // bin by bin; inside a bin by (group, x, y), so that consecutive items of
    // one group share what the kernel keeps in LDS
    std::vector<size_t> idx(n);
    for (size_t i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
        if (bin_of[a] != bin_of[b]) return bin_of[a] < bin_of[b];
        const int4 &p = items[a], &q = items[b];
        if (p.x != q.x) return p.x < q.x;
        if (p.y != q.y) return p.y < q.y;
        return p.z < q.z;
    });
    dealt.resize(n);
    bounds.assign(nbins + 1, 0);
    for (size_t i = 0; i < n; ++i) {
        dealt[i] = items[idx[i]];
        ++bounds[bin_of[idx[i]] + 1];
    }
    for (int b = 0; b < nbins; ++b) bounds[b + 1] += bounds[b];
    pl.cost_max = 0;
    double sum = 0;
    for (double l : load) {
        pl.cost_max = std::max(pl.cost_max, l);
        sum += l;
    }
    pl.cost_mean = sum / nbins;
    pl.bin_cost = load;
}

void* upload_owned(PersistentLaunch& pl, const void* data, size_t bytes) {
    void* d = dmalloc(std::max<size_t>(bytes, 16), "persistent work list");
    pl.buffers.push_back(d);
    if (bytes) check_hip(hipMemcpy(d, data, bytes, hipMemcpyHostToDevice), "persistent work list upload");
    return d;
}

int resident_blocks(const void* kernel, int threads, size_t lds) {
    int dev = 0, cus = 0, per = 0;
    check_hip(hipGetDevice(&dev), "device");
    check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, lds), "occupancy");
    return std::max(1, cus * std::max(1, per));
}

void dfree(void* p) noexcept {
    if (!p) return;
    HookedBlock b{};
    bool hooked = false;
    {
        std::lock_guard<std::mutex> lock(g_alloc_mutex);
        auto it = g_hooked.find(p);
        if (it != g_hooked.end()) {
            b = it->second;
            g_hooked.erase(it);
            hooked = true;
        }
    }
    if (!hooked) {
        (void)hipFree(p);  // synchronizes with the device itself
        return;
    }
    (void)hipDeviceSynchronize();
    b.by.free(p, b.bytes, b.by.user);
}

namespace {

thread_local std::string g_last_error;

zgl_status fail(zgl_status s, const std::string& msg) {
    g_last_error = msg;
    return s;
}

// A pointer to T that is not aligned to sizeof(T) is no valid T* (and the
// kernels' loads assume element alignment).
template <class... P>
bool misaligned(const P*... ptrs) {
    return ((reinterpret_cast<uintptr_t>(ptrs) % sizeof(P) != 0) || ...);
}
constexpr const char* kMisaligned = "pointer not aligned to its element size";

constexpr size_t kMaxOrder = 4096;

size_t env_mb(const char* name, size_t dflt) {
    const char* v = std::getenv(name);
    if (!v || !*v) return dflt;
    char* end = nullptr;
    const unsigned long long x = std::strtoull(v, &end, 10);
    return (end && *end == 0 && x > 0) ? (size_t)x : dflt;
}

}  // namespace
}  // namespace zgl

namespace zgl {
// Device scratch of one pipeline: the spectrum and flm of a chunk of
// instances, and the refinement residual grids / corrections.
struct Scratch {
    void* spec = nullptr;
    size_t spec_cap = 0;  // bytes
    void* flm = nullptr;
    size_t flm_cap = 0;   // bytes
    double* rgrid = nullptr;  // refinement residual grids (chunk)
    size_t rgrid_cap = 0;     // doubles
    double* rcoef = nullptr;  // refinement corrections / scaled coefficients (chunk)
    size_t rcoef_cap = 0;     // doubles
    void release() {
        for (void* q : {spec, flm, (void*)rgrid, (void*)rcoef}) dfree(q);
        *this = Scratch{};
    }
};
}  // namespace zgl

struct zgl_plan {
    zgl_kind kind;
    int znorm, shnorm, phase, layout;
    zgl_precision precision;
    size_t order = 0;
    std::map<size_t, zgl::OrderTables> tables;
    zgl::Scratch main;                // the batched entry points
    double* dgrid = nullptr;
    size_t dgrid_cap = 0;  // doubles
    double* dcoef = nullptr;
    size_t dcoef_cap = 0;  // doubles
    // autotuning (zgl_plan_autotune): (order, bytes per real of the
    // arithmetic, stage, bytes per grid value) -> chosen FFT tile width or
    // instances per workgroup
    bool autotune = false;
    std::map<std::tuple<size_t, int, int, int>, int> tuned;
    // instances per chunk at most (0: no limit beyond the scratch budget);
    // set by zgl_forward_noalloc for the duration of its call
    size_t chunk_cap = 0;
    // stage timing (zgl_profile_enable)
    bool profiling = false;
    struct Rec { int stage; int nb; hipEvent_t start, stop; };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    // ragged batches (zgl_forward_ragged / zgl_backward_ragged)
    struct Ragged {
        std::vector<hipStream_t> lanes;   // internal streams, default priority
        std::vector<zgl::Scratch> scratch;  // one per lane
        std::vector<hipEvent_t> ev;       // fork, then one join event per lane
        // per direction (0 backward, 1 forward): the last job list and its
        // captured pipeline; the graphs hold the lane scratch addresses
        struct Cached {
            std::vector<zgl_job> jobs;
            bool valid = false;
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            long* offs = nullptr;  // device: grid offsets [njobs], coefficient offsets [njobs]
            size_t offs_cap = 0;   // longs
            // recorded on the caller's stream after each launch that reads
            // `offs` or runs `exec`: both are rewritten or destroyed only once
            // that work has completed
            hipEvent_t done = nullptr;
            // the persistent pipeline of the job list (valid && persistent):
            // its launches in order, wave by wave, on `pscratch`
            bool persistent = false;
            std::vector<zgl::PersistentLaunch> plist;
        } cache[2];
        std::vector<zgl::Scratch> pscratch;  // per lane: spectra and flm of one wave of the persistent pipeline
        std::vector<double> lane_cost;  // modelled work per lane of the last build
        bool grid16 = false;  // every grid of the job list being built is 16-byte aligned
    } rg;
    // zgl_evaluate: tables and staging at the last evaluation grid
    zgl::EvalCache eval;
    double* dpts = nullptr;  // host-memory calls: the points, uploaded
    size_t dpts_cap = 0;
    // rotations: the Wigner d(pi/2) table of degrees < wtab_order, work lists
    // per (kind, order, batch), the intermediate and the uploaded angles
    double* wtab = nullptr;
    int wtab_order = 0;
    std::map<std::tuple<int, int, int, int>, std::pair<int4*, int>> rot_work;  // (fused?, kind, order, batch)
    double* rtemp = nullptr;
    size_t rtemp_cap = 0;
    double* dang = nullptr;
    size_t dang_cap = 0;
};

namespace zgl {
namespace {

// Set while zgl_forward_noalloc runs on this thread: a buffer that would
// have to grow, or a table not yet built, is an error there, never an
// allocation.
thread_local bool t_no_alloc = false;

size_t grid_size(const zgl_plan* p, size_t order) {
    if (order == 0) return 0;
    const size_t lon_min = 2 * order - 1;
    if (p->kind == ZGL_ZERNIKE) return lon_min * (order + 1) * ((order + 1 + 7) / 8 * 8);
    if (p->layout == ZGL_LAYOUT_LATLON) return order * ((lon_min + 7) / 8 * 8);
    return lon_min * ((order + 7) / 8 * 8);
}

size_t coeff_size(const zgl_plan* p, size_t order) {
    if (p->kind == ZGL_ZERNIKE) return (order + 1) * (order + 3) * (2 * order + 1) / 24;
    return order * (order + 1) / 2;
}

const OrderTables& tables_for(zgl_plan* p, size_t order) {
    auto it = p->tables.find(order);
    if (it != p->tables.end()) return it->second;
    if (t_no_alloc)
        throw Error{ZGL_ERROR_INVALID_VALUE,
                    "no tables for order " + std::to_string(order) +
                        ": create or resize the plan to that order before the no-allocation call"};
    OrderTables t;
    try {
        build_tables(t, p->kind, p->znorm, p->shnorm, p->phase, p->layout, (int)order,
                     p->precision != ZGL_FP64, nullptr);
    } catch (...) {
        free_tables(t);
        throw;
    }
    return p->tables.emplace(order, t).first->second;
}

void no_alloc_check(size_t need, size_t cap) {
    if (t_no_alloc && need > cap)
        throw Error{ZGL_ERROR_INVALID_VALUE, "scratch of the no-allocation path too small (internal)"};
}

// Growing a buffer frees the old one: wait for queued work that may use it.
template <class T>
void ensure(T*& ptr, size_t& cap, size_t need) {
    if (need <= cap) return;
    no_alloc_check(need, cap);
    if (ptr) check_hip(hipDeviceSynchronize(), "scratch growth");
    dfree(ptr);
    ptr = nullptr;
    cap = 0;
    ptr = dmalloc_n<T>(need, "scratch allocation");
    cap = need;
}

void ragged_release(zgl_plan::Ragged& r) {
    for (auto& c : r.cache) {
        if (c.done) (void)hipEventSynchronize(c.done);
        if (c.exec) (void)hipGraphExecDestroy(c.exec);
        if (c.graph) (void)hipGraphDestroy(c.graph);
        dfree(c.offs);
        if (c.done) (void)hipEventDestroy(c.done);
    }
    for (auto l : r.lanes)
        if (l) (void)hipStreamDestroy(l);
    for (auto e : r.ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& sc : r.scratch) sc.release();
    for (auto& c : r.cache) release(c.plist);
    for (auto& sc : r.pscratch) sc.release();
    r = zgl_plan::Ragged{};
}

void ensure_bytes(void*& ptr, size_t& cap, size_t need) {
    if (need <= cap) return;
    no_alloc_check(need, cap);
    if (ptr) check_hip(hipDeviceSynchronize(), "scratch growth");
    dfree(ptr);
    ptr = nullptr;
    cap = 0;
    ptr = dmalloc(need, "scratch allocation");
    cap = need;
}

struct Chunking {
    int nb;
    long spec_stride;  // double2
    long flm_stride;   // double2
};

Chunking chunking(const zgl_plan* p, const Geom& g, size_t batch, size_t csize) {
    Chunking c;
    c.spec_stride = (long)g.nm * 2 * g.jh * g.ncol;
    // Zernike: flm; SH: the coefficient-major staging of the Legendre stage
    c.flm_stride = (long)g.order * (g.order + 1) / 2 * (p->kind == ZGL_ZERNIKE ? g.ncol : 1);
    const size_t per = (size_t)(c.spec_stride + c.flm_stride) * csize;
    size_t budget, cap;
    if (p->kind == ZGL_ZERNIKE) {
        // Large chunks keep every CU busy through each stage and cut launches;
        // measured faster than chunks sized to stay in the 256 MB Infinity
        // Cache (order 128: 64 instances of 27 MB FP32 / 54 MB FP64 scratch).
        // The cap keeps the work per chunk about that of 64 instances at order
        // 128 (memory ~ N^3): small orders take far more instances per chunk
        // (order 16: 608k -> 1.74M transforms/s with 1024 instead of 64).
        budget = env_mb("ZGL_SCRATCH_MB", 4096) << 20;
        const double scale = 128.0 / (double)std::max<size_t>(1, g.order);
        cap = (size_t)std::min(8192.0, std::max(64.0, 64.0 * scale * scale * scale));
    } else {
        // Instances form the GEMM column axis: larger chunks reuse the P table more.
        budget = env_mb("ZGL_SH_SCRATCH_MB", 4096) << 20;
        cap = 1024;
    }
    size_t nb = per ? budget / per : cap;
    nb = std::max<size_t>(1, std::min({nb, cap, batch}));
    if (p->chunk_cap) nb = std::min(nb, p->chunk_cap);
    c.nb = (int)nb;
    return c;
}

enum Stage { ST_FFT_FWD, ST_LEG_FWD, ST_RAD_FWD, ST_RAD_BWD, ST_LEG_BWD, ST_FFT_BWD,
             ST_EVAL_TABLES, ST_EVAL_RAD, ST_EVAL_LEG, ST_EVAL_LON, ST_ROTATE, ST_ROTATE_WIDE, ST_POLAR,
             ST_SPECTRUM, ST_FUSED_FWD, ST_FUSED_BWD, ST_COUNT };
const char* const kStageNames[ST_COUNT] = {"fft_forward", "legendre_forward", "radial_forward",
                                           "radial_backward", "legendre_backward", "fft_backward",
                                           "eval_tables", "eval_radial", "eval_legendre",
                                           "eval_longitude", "rotate", "rotate_wide", "polar_rotate",
                                           "power_spectrum", "legendre_radial_forward",
                                           "radial_legendre_backward"};

hipEvent_t pooled_event(zgl_plan* p) {
    if (!p->pool.empty()) {
        hipEvent_t e = p->pool.back();
        p->pool.pop_back();
        return e;
    }
    hipEvent_t e;
    check_hip(hipEventCreate(&e), "event create");
    return e;
}

// Runs `launch` between two events on stream s when the plan is profiling.
template <class F>
void timed(zgl_plan* p, Stage st, int nb, hipStream_t s, F&& launch) {
    if (!p->profiling) {
        launch();
        return;
    }
    zgl_plan::Rec r{st, nb, pooled_event(p), pooled_event(p)};
    check_hip(hipEventRecord(r.start, s), "event record");
    launch();
    check_hip(hipEventRecord(r.stop, s), "event record");
    p->recs.push_back(r);
}

double shnorm_value(int sh_norm) { return sh_norm == ZGL_SHNORM_QM ? 1.0 : 1.0 / (4.0 * M_PI); }

// The forward transform's prefactor on the FFT: shnorm * 2 pi / Nphi, halved
// for Zernike (include/zernike_glq_transformer.hpp:710-720,
// include/sh_glq_transformer.hpp:757-765).
double forward_prefactor(const zgl_plan* p, const Geom& g) {
    const double twopi_n = 2.0 * M_PI / (double)g.nphi;
    return p->kind == ZGL_ZERNIKE ? shnorm_value(p->shnorm) * 0.5 * twopi_n
                                  : shnorm_value(p->shnorm) * twopi_n;
}

// ---------------------------------------------------------------- autotuner
// The first call of a stage at a given order and arithmetic, with autotuning
// on, times every candidate (one untimed run, one timed with events) on the
// call's own operands and keeps the fastest. The stage kernels only read
// their input and overwrite their output, so running one several times leaves
// the same result. Nothing is tuned while the stream is being captured.
template <class F>
int tune(hipStream_t s, const std::vector<int>& cands, F&& launch) {
    hipEvent_t a, b;
    check_hip(hipEventCreate(&a), "tune event");
    check_hip(hipEventCreate(&b), "tune event");
    // one warm-up launch each, then three rounds over all candidates (so a
    // drift of the device clock or of the caches affects every candidate
    // alike); a candidate's time is its best of the three
    for (int c : cands) launch(c);
    std::vector<float> best(cands.size(), std::numeric_limits<float>::max());
    for (int rep = 0; rep < 3; ++rep)
        for (size_t i = 0; i < cands.size(); ++i) {
            check_hip(hipEventRecord(a, s), "tune");
            launch(cands[i]);
            check_hip(hipEventRecord(b, s), "tune");
            check_hip(hipEventSynchronize(b), "tune");
            float ms = 0;
            check_hip(hipEventElapsedTime(&ms, a, b), "tune");
            best[i] = std::min(best[i], ms);
        }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    // the list is in order of preference: a later candidate has to be faster
    // by more than 1% (the launch-to-launch noise) to be picked
    size_t pick = 0;
    for (size_t i = 1; i < cands.size(); ++i)
        if (best[i] < 0.99f * best[pick]) pick = i;
    return cands[pick];
}

// `cands` yields the candidates (built only when a stage is tuned)
template <class CF, class F>
int tuned_choice(zgl_plan* p, size_t order, int rbytes, int gbytes, int stage, hipStream_t s,
                 CF&& cands_of, F&& launch) {
    const auto key = std::make_tuple(order, rbytes, stage, gbytes);
    auto it = p->tuned.find(key);
    if (it != p->tuned.end()) return it->second;
    if (!p->autotune) return 0;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return 0;
    const std::vector<int> cands = cands_of();
    if (cands.size() < 2) return 0;
    const int c = tune(s, cands, launch);
    p->tuned[key] = c;
    return c;
}

// the untuned default first; kGroupTiles | t: workgroups of 4 t column tiles
// of the flattened (instance, column) index (every wave the same count of
// tiles, where G instances per workgroup leave a partial tile at order 128)
const std::vector<int> kGroupCandidates = {4, 2, 8, 1, 3, 6, 12, kGroupTiles | 2, kGroupTiles | 3,
                                           kGroupTiles | 4, kGroupTiles | 6};
// the radial stages also take 16 instances per workgroup (no flag bits there)
const std::vector<int> kRadialCandidates = {4, 2, 8, 1, 16};
// BF16x3 Legendre products (kGroupX3) at every group size: the candidates of
// ZGL_BF16X3 plans. The FP32 modes keep the FP32 MFMA products: timed alone,
// a BF16x3 backward stage at order 128 looked faster than the FP32 one, but
// inside the pipeline it ran slower (468 against 410-425 us per 64
// instances), so the two are not ranked by the first-use timing.
const std::vector<int> kLegendreCandidatesX3 = {4 | kGroupX3, 2 | kGroupX3, 8 | kGroupX3, 1 | kGroupX3};

template <class T>
const std::vector<int>& legendre_candidates(const zgl_plan* p) {
    return sizeof(T) == 4 && p->precision == ZGL_BF16X3 ? kLegendreCandidatesX3 : kGroupCandidates;
}

template <class T>
int legendre_math(const zgl_plan* p) {
    return sizeof(T) == 4 && p->precision == ZGL_BF16X3 ? kGroupX3 : 0;
}

// The fused FP32 colatitude + radial kernels replace the two stage kernels
// for FP32 arithmetic at orders up to 128 (not for the BF16x3 mode, whose
// colatitude products are formed differently).
template <class T>
bool use_fused(const zgl_plan* p, const OrderTables& t) {
    return sizeof(T) == 4 && p->precision != ZGL_BF16X3 && t.rfuse_f && zt_fused_fits(t);
}

// T: arithmetic (double or float); G: grid element type in memory.
// adjoint: the transpose of the backward transform (no quadrature weights,
// no prefactor, no kappa_n), otherwise the forward transform.
// Whether every instance base of x is 16-byte aligned (for a ragged batch:
// the grids of the job list being run).
template <class P>
bool aligned16(const zgl_plan* p, const Inst<P>& x) {
    if (x.offs) return p->rg.grid16;
    return reinterpret_cast<uintptr_t>(x.base) % 16 == 0 && (x.stride * (long)sizeof(P)) % 16 == 0;
}

template <class T, class G>
void run_forward(zgl_plan* p, Scratch& sc, Inst<const G> grid, size_t N, Inst<double> coeffs,
                 size_t nt, size_t batch, hipStream_t s, bool adjoint = false) {
    using C = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
    const OrderTables& t = tables_for(p, N);
    const Geom& g = t.g;
    const Chunking ch = chunking(p, g, batch, sizeof(C));
    ensure_bytes(sc.spec, sc.spec_cap, (size_t)ch.spec_stride * ch.nb * sizeof(C));
    if (ch.flm_stride) ensure_bytes(sc.flm, sc.flm_cap, (size_t)ch.flm_stride * ch.nb * sizeof(C));
    FftLaunch L;
    L.g = &g;
    L.t = &t;
    L.layout_kind = (p->kind == ZGL_SH && p->layout == ZGL_LAYOUT_LATLON) ? GRID_ROWS : GRID_COLS;
    L.grid_aligned16 = aligned16(p, grid);
    L.nm_used = (int)nt;
    L.weighted = !adjoint;
    L.scale = adjoint ? 1.0 : forward_prefactor(p, g);
    L.center_factor = (adjoint || p->kind == ZGL_ZERNIKE) ? 1.0 : 2.0;
    const int kappa = !adjoint && p->znorm == ZGL_ZNORM_UNNORMED ? 1 : 0;
    for (size_t b0 = 0; b0 < batch; b0 += ch.nb) {
        const int nb = (int)std::min<size_t>(ch.nb, batch - b0);
        L.tile = tuned_choice(p, N, (int)sizeof(T), (int)sizeof(G), ST_FFT_FWD, s, [&] { return fft_tiles(t.fft, sizeof(T) == 4); },
                              [&](int c) {
                                  FftLaunch L2 = L;
                                  L2.tile = c;
                                  fft_forward<T, G>(L2, grid.from(b0), sc.spec, ch.spec_stride, nb, s);
                              });
        timed(p, ST_FFT_FWD, nb, s, [&] { fft_forward<T, G>(L, grid.from(b0), sc.spec, ch.spec_stride, nb, s); });
        if (p->kind == ZGL_ZERNIKE && use_fused<T>(p, t)) {
            ZtGemmArgs a{&t, (int)nt, kappa, sc.spec, ch.spec_stride, sc.flm, ch.flm_stride,
                         coeffs.from(b0), nb};
            timed(p, ST_FUSED_FWD, nb, s, [&] { zt_fused_forward(a, s); });
        } else if (p->kind == ZGL_ZERNIKE) {
            ZtGemmArgs a{&t, (int)nt, kappa, sc.spec, ch.spec_stride, sc.flm, ch.flm_stride,
                         coeffs.from(b0), nb};
            a.group = tuned_choice(p, N, (int)sizeof(T), (int)sizeof(G), ST_LEG_FWD, s, [&] { return legendre_candidates<T>(p); }, [&](int c) {
                ZtGemmArgs b = a;
                b.group = c;
                zt_legendre_forward<T>(b, s);
            });
            a.group |= legendre_math<T>(p);
            timed(p, ST_LEG_FWD, nb, s, [&] { zt_legendre_forward<T>(a, s); });
            a.group = tuned_choice(p, N, (int)sizeof(T), (int)sizeof(G), ST_RAD_FWD, s, [&] { return kRadialCandidates; }, [&](int c) {
                ZtGemmArgs b = a;
                b.group = c;
                zt_radial_forward<T>(b, s);
            });
            timed(p, ST_RAD_FWD, nb, s, [&] { zt_radial_forward<T>(a, s); });
        } else {
            StGemmArgs a{&t, (int)nt, sc.spec, ch.spec_stride, coeffs.from(b0), nb, sc.flm};
            timed(p, ST_LEG_FWD, nb, s, [&] { st_legendre_forward<T>(a, s); });
        }
    }
}

// adjoint: the transpose of the forward transform, W * B(prefactor * kappa_n * c)
// (the prefactor and the quadrature weights W applied to the output grid,
// kappa_n to a scaled copy of the coefficients), otherwise the backward
// transform.
template <class T, class G>
void run_backward(zgl_plan* p, Scratch& sc, Inst<const double> coeffs, size_t nt, Inst<G> grid,
                  size_t N, size_t batch, hipStream_t s, bool adjoint = false) {
    using C = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
    const OrderTables& t = tables_for(p, N);
    const Geom& g = t.g;
    const Chunking ch = chunking(p, g, batch, sizeof(C));
    ensure_bytes(sc.spec, sc.spec_cap, (size_t)ch.spec_stride * ch.nb * sizeof(C));
    if (ch.flm_stride) ensure_bytes(sc.flm, sc.flm_cap, (size_t)ch.flm_stride * ch.nb * sizeof(C));
    FftLaunch L;
    L.g = &g;
    L.t = &t;
    L.layout_kind = (p->kind == ZGL_SH && p->layout == ZGL_LAYOUT_LATLON) ? GRID_ROWS : GRID_COLS;
    L.grid_aligned16 = aligned16(p, grid);
    L.nm_used = (int)nt;
    L.weighted = adjoint;
    L.scale = adjoint ? forward_prefactor(p, g) : 1.0;
    L.center_factor = (adjoint && p->kind == ZGL_SH) ? 2.0 : 1.0;
    const bool kappa = adjoint && p->kind == ZGL_ZERNIKE && p->znorm == ZGL_ZNORM_UNNORMED;
    const long ncw = 2 * (long)coeff_size(p, nt);
    if (kappa) ensure(sc.rcoef, sc.rcoef_cap, (size_t)ncw * ch.nb);
    for (size_t b0 = 0; b0 < batch; b0 += ch.nb) {
        const int nb = (int)std::min<size_t>(ch.nb, batch - b0);
        // the kernels read coefficients through Inst<double>; they never write them
        Inst<double> cin{const_cast<double*>(coeffs.base), coeffs.stride, coeffs.offs};
        cin = cin.from(b0);
        if (kappa) {
            coeff_kappa_scale(cin, sc.rcoef, ncw, (int)nt, 1, nb, s);
            cin = Inst<double>{sc.rcoef, ncw, nullptr};
        }
        if (p->kind == ZGL_ZERNIKE && use_fused<T>(p, t)) {
            ZtGemmArgs a{&t, (int)nt, 0, sc.spec, ch.spec_stride, sc.flm, ch.flm_stride, cin, nb};
            timed(p, ST_FUSED_BWD, nb, s, [&] { zt_fused_backward(a, s); });
        } else if (p->kind == ZGL_ZERNIKE) {
            ZtGemmArgs a{&t, (int)nt, 0, sc.spec, ch.spec_stride, sc.flm, ch.flm_stride, cin, nb};
            a.group = tuned_choice(p, N, (int)sizeof(T), (int)sizeof(G), ST_RAD_BWD, s, [&] { return kRadialCandidates; }, [&](int c) {
                ZtGemmArgs b = a;
                b.group = c;
                zt_radial_backward<T>(b, s);
            });
            timed(p, ST_RAD_BWD, nb, s, [&] { zt_radial_backward<T>(a, s); });
            a.group = tuned_choice(p, N, (int)sizeof(T), (int)sizeof(G), ST_LEG_BWD, s, [&] { return legendre_candidates<T>(p); }, [&](int c) {
                ZtGemmArgs b = a;
                b.group = c;
                zt_legendre_backward<T>(b, s);
            });
            a.group |= legendre_math<T>(p);
            timed(p, ST_LEG_BWD, nb, s, [&] { zt_legendre_backward<T>(a, s); });
        } else {
            StGemmArgs a{&t, (int)nt, sc.spec, ch.spec_stride, cin, nb, sc.flm};
            timed(p, ST_LEG_BWD, nb, s, [&] { st_legendre_backward<T>(a, s); });
        }
        L.tile = tuned_choice(p, N, (int)sizeof(T), (int)sizeof(G), ST_FFT_BWD, s, [&] { return fft_tiles(t.fft, sizeof(T) == 4); },
                              [&](int c) {
                                  FftLaunch L2 = L;
                                  L2.tile = c;
                                  fft_backward<T, G>(L2, sc.spec, ch.spec_stride, grid.from(b0), nb, s);
                              });
        timed(p, ST_FFT_BWD, nb, s, [&] { fft_backward<T, G>(L, sc.spec, ch.spec_stride, grid.from(b0), nb, s); });
    }
}

// Forward in the plan's precision. With ZGL_FP32_REFINE each chunk of
// instances is refined once: c += F32(g - B64(c)), the residual in FP64.
template <class G>
void forward_any(zgl_plan* p, Scratch& sc, Inst<const G> grid, size_t N, Inst<double> coeffs,
                 size_t nt, size_t batch, hipStream_t s) {
    if (p->precision == ZGL_FP64) {
        if constexpr (std::is_same_v<G, double>) run_forward<double, G>(p, sc, grid, N, coeffs, nt, batch, s);
        return;
    }
    if (p->precision == ZGL_FP32 || p->precision == ZGL_BF16X3) {
        run_forward<float, G>(p, sc, grid, N, coeffs, nt, batch, s);
        return;
    }
    const long gsz = (long)grid_size(p, N);
    const long ncw = 2 * (long)coeff_size(p, nt);
    const Chunking ch = chunking(p, tables_for(p, N).g, batch, sizeof(float2));
    const size_t nb_max = (size_t)ch.nb;
    ensure(sc.rgrid, sc.rgrid_cap, (size_t)gsz * nb_max);
    ensure(sc.rcoef, sc.rcoef_cap, (size_t)ncw * nb_max);
    const Inst<double> rg{sc.rgrid, gsz, nullptr};
    const Inst<double> rc{sc.rcoef, ncw, nullptr};
    for (size_t b0 = 0; b0 < batch; b0 += nb_max) {
        const int nb = (int)std::min(nb_max, batch - b0);
        const Inst<const G> g = grid.from(b0);
        const Inst<double> c = coeffs.from(b0);
        run_forward<float, G>(p, sc, g, N, c, nt, nb, s);
        run_backward<double, double>(p, sc, c, nt, rg, N, nb, s);
        grid_residual<G>(g, sc.rgrid, gsz, gsz, nb, s);
        run_forward<float, double>(p, sc, rg, N, rc, nt, nb, s);
        coeff_accumulate(c, sc.rcoef, ncw, ncw, nb, s);
    }
}

template <class G>
void backward_any(zgl_plan* p, Scratch& sc, Inst<const double> coeffs, size_t nt, Inst<G> grid,
                  size_t N, size_t batch, hipStream_t s) {
    if (p->precision == ZGL_FP64) {
        if constexpr (std::is_same_v<G, double>) run_backward<double, G>(p, sc, coeffs, nt, grid, N, batch, s);
    } else {
        run_backward<float, G>(p, sc, coeffs, nt, grid, N, batch, s);
    }
}

// The two transposes, in the plan's arithmetic (FP64, or FP32 for both FP32
// modes: refinement does not apply to a transpose).
void backward_adjoint_any(zgl_plan* p, Scratch& sc, Inst<const double> grid, size_t N,
                          Inst<double> coeffs, size_t nt, size_t batch, hipStream_t s) {
    if (p->precision == ZGL_FP64) run_forward<double, double>(p, sc, grid, N, coeffs, nt, batch, s, true);
    else run_forward<float, double>(p, sc, grid, N, coeffs, nt, batch, s, true);
}

void forward_adjoint_any(zgl_plan* p, Scratch& sc, Inst<const double> coeffs, size_t nt,
                         Inst<double> grid, size_t N, size_t batch, hipStream_t s) {
    if (p->precision == ZGL_FP64) run_backward<double, double>(p, sc, coeffs, nt, grid, N, batch, s, true);
    else run_backward<float, double>(p, sc, coeffs, nt, grid, N, batch, s, true);
}

template <class F>
zgl_status guarded(F&& f) {
    try {
        return f();
    } catch (const Error& e) {
        return fail(e.status, e.what);
    } catch (const std::bad_alloc&) {
        return fail(ZGL_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (...) {
        return fail(ZGL_ERROR_HIP, "unexpected exception");
    }
}

// Resolves the device address of the no-allocation path's static scratch on
// the current device, when there is one (defined with the array).
void resolve_static_scratch();

zgl_status check_conventions(zgl_kind kind, int zn, int sn, int sp, int layout) {
    if (kind != ZGL_ZERNIKE && kind != ZGL_SH) return fail(ZGL_ERROR_INVALID_VALUE, "unknown transform kind");
    if (zn != ZGL_ZNORM_NORMED && zn != ZGL_ZNORM_UNNORMED)
        return fail(ZGL_ERROR_INVALID_VALUE, "zernike_norm must be 0 (normed) or 1 (unnormed)");
    if (sn != ZGL_SHNORM_GEO && sn != ZGL_SHNORM_QM)
        return fail(ZGL_ERROR_INVALID_VALUE, "sh_norm must be 0 (geo) or 1 (qm)");
    if (sp != ZGL_PHASE_NONE && sp != ZGL_PHASE_CS)
        return fail(ZGL_ERROR_INVALID_VALUE, "sh_phase must be -1 (none) or +1 (cs)");
    if (kind == ZGL_SH && layout != ZGL_LAYOUT_LONLAT && layout != ZGL_LAYOUT_LATLON)
        return fail(ZGL_ERROR_INVALID_VALUE, "layout must be 0 (LonLat) or 1 (LatLon)");
    return ZGL_SUCCESS;
}

}  // namespace
}  // namespace zgl

using namespace zgl;

extern "C" {

zgl_status zgl_plan_create(zgl_plan** plan, zgl_kind kind, int zernike_norm, int sh_norm,
                           int sh_phase, int layout, size_t order, zgl_precision precision) {
    if (!plan) return fail(ZGL_ERROR_INVALID_VALUE, "plan pointer is NULL");
    *plan = nullptr;
    zgl_status st = check_conventions(kind, zernike_norm, sh_norm, sh_phase, layout);
    if (st != ZGL_SUCCESS) return st;
    if (precision != ZGL_FP64 && precision != ZGL_FP32 && precision != ZGL_FP32_REFINE &&
        precision != ZGL_BF16X3)
        return fail(ZGL_ERROR_NOT_SUPPORTED, "unsupported precision");
    if (order > kMaxOrder) return fail(ZGL_ERROR_INVALID_VALUE, "order exceeds the supported maximum");
    return guarded([&]() -> zgl_status {
        zgl_plan* p = new zgl_plan;
        p->kind = kind;
        p->znorm = zernike_norm;
        p->shnorm = sh_norm;
        p->phase = sh_phase;
        p->layout = kind == ZGL_SH ? layout : 0;
        p->precision = precision;
        p->order = order;
        {
            const char* e = std::getenv("ZGL_AUTOTUNE");
            p->autotune = e && e[0] == '1';
        }
        try {
            if (order) tables_for(p, order);
        } catch (...) {
            zgl_plan_destroy(p);
            throw;
        }
        resolve_static_scratch();
        *plan = p;
        return ZGL_SUCCESS;
    });
}

zgl_status zgl_plan_destroy(zgl_plan* plan) {
    if (!plan) return ZGL_SUCCESS;
    (void)hipDeviceSynchronize();
    for (auto& kv : plan->tables) free_tables(kv.second);
    for (auto& r : plan->recs) {
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    for (auto e : plan->pool) (void)hipEventDestroy(e);
    plan->main.release();
    ragged_release(plan->rg);
    plan->eval.release();
    for (auto& kv : plan->rot_work) dfree(kv.second.first);
    for (void* q : {(void*)plan->dgrid, (void*)plan->dcoef, (void*)plan->dpts, (void*)plan->wtab,
                    (void*)plan->rtemp, (void*)plan->dang})
        dfree(q);
    delete plan;
    return ZGL_SUCCESS;
}

zgl_status zgl_plan_resize(zgl_plan* plan, size_t order) {
    if (!plan) return fail(ZGL_ERROR_INVALID_VALUE, "plan is NULL");
    if (order > kMaxOrder) return fail(ZGL_ERROR_INVALID_VALUE, "order exceeds the supported maximum");
    return guarded([&]() -> zgl_status {
        if (order) tables_for(plan, order);
        plan->order = order;
        return ZGL_SUCCESS;
    });
}

size_t zgl_plan_order(const zgl_plan* plan) { return plan ? plan->order : 0; }

size_t zgl_grid_size(const zgl_plan* plan, size_t order) { return plan ? grid_size(plan, order) : 0; }

size_t zgl_coeff_size(const zgl_plan* plan, size_t order) { return plan ? coeff_size(plan, order) : 0; }

}  // extern "C"

namespace zgl {
namespace {

// grid -> coefficients: the forward transform, or (adjoint) the transpose of
// the backward transform
template <class G>
zgl_status forward_entry(zgl_plan* plan, const G* grid, size_t grid_order, double* coeffs,
                         size_t exp_order, size_t batch, ptrdiff_t grid_stride,
                         ptrdiff_t coeff_stride, zgl_mem mem, void* stream, bool adjoint = false) {
    if (!plan) return fail(ZGL_ERROR_INVALID_VALUE, "plan is NULL");
    if (grid_order > kMaxOrder) return fail(ZGL_ERROR_INVALID_VALUE, "grid order exceeds the supported maximum");
    if (mem != ZGL_MEM_DEVICE && mem != ZGL_MEM_HOST) return fail(ZGL_ERROR_INVALID_VALUE, "unknown memory kind");
    if (std::is_same_v<G, float> && plan->precision == ZGL_FP64)
        return fail(ZGL_ERROR_NOT_SUPPORTED, "FP32 grids need an FP32 plan");
    const size_t nt = std::min(exp_order, grid_order);
    const long gsz = (long)grid_size(plan, grid_order);
    const long csz = 2 * (long)coeff_size(plan, exp_order);
    const long gs = grid_stride ? (long)grid_stride : gsz;
    const long cs = coeff_stride ? (long)coeff_stride : csz;
    if (batch > 1 && (gs < gsz || cs < 2 * (long)coeff_size(plan, nt)))
        return fail(ZGL_ERROR_INVALID_VALUE, "instance strides overlap");
    return guarded([&]() -> zgl_status {
        if (grid_order) tables_for(plan, grid_order);
        plan->order = grid_order;
        if (batch == 0 || nt == 0) return ZGL_SUCCESS;
        if (!grid || !coeffs) return fail(ZGL_ERROR_INVALID_VALUE, "NULL grid or coefficient pointer");
        if (misaligned(grid, coeffs)) return fail(ZGL_ERROR_INVALID_VALUE, kMisaligned);
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (mem == ZGL_MEM_DEVICE) {
            const Inst<const G> gi{grid, gs, nullptr};
            const Inst<double> ci{coeffs, cs, nullptr};
            if constexpr (std::is_same_v<G, double>) {
                if (adjoint) backward_adjoint_any(plan, plan->main, gi, grid_order, ci, nt, batch, s);
                else forward_any<G>(plan, plan->main, gi, grid_order, ci, nt, batch, s);
            } else {
                forward_any<G>(plan, plan->main, gi, grid_order, ci, nt, batch, s);
            }
            check_hip(hipGetLastError(), "forward");
            return ZGL_SUCCESS;
        }
        const size_t ncw = 2 * coeff_size(plan, nt);  // doubles written per instance
        ensure(plan->dgrid, plan->dgrid_cap, ((size_t)gsz * batch * sizeof(G) + 7) / 8);
        ensure(plan->dcoef, plan->dcoef_cap, ncw * batch);
        G* dg = reinterpret_cast<G*>(plan->dgrid);
        check_hip(hipMemcpy2DAsync(dg, gsz * sizeof(G), grid, gs * sizeof(G), gsz * sizeof(G), batch,
                                   hipMemcpyHostToDevice, s),
                  "grid upload");
        const Inst<const G> dgi{dg, gsz, nullptr};
        const Inst<double> dci{plan->dcoef, (long)ncw, nullptr};
        if constexpr (std::is_same_v<G, double>) {
            if (adjoint) backward_adjoint_any(plan, plan->main, dgi, grid_order, dci, nt, batch, s);
            else forward_any<G>(plan, plan->main, dgi, grid_order, dci, nt, batch, s);
        } else {
            forward_any<G>(plan, plan->main, dgi, grid_order, dci, nt, batch, s);
        }
        check_hip(hipMemcpy2DAsync(coeffs, cs * sizeof(double), plan->dcoef, ncw * sizeof(double),
                                   ncw * sizeof(double), batch, hipMemcpyDeviceToHost, s),
                  "coefficient download");
        check_hip(hipStreamSynchronize(s), "forward");
        return ZGL_SUCCESS;
    });
}

// coefficients -> grid: the backward transform, or (adjoint) the transpose of
// the forward transform
template <class G>
zgl_status backward_entry(zgl_plan* plan, const double* coeffs, size_t exp_order, G* grid,
                          size_t grid_order, size_t batch, ptrdiff_t coeff_stride,
                          ptrdiff_t grid_stride, zgl_mem mem, void* stream, bool adjoint = false) {
    if (!plan) return fail(ZGL_ERROR_INVALID_VALUE, "plan is NULL");
    if (grid_order > kMaxOrder) return fail(ZGL_ERROR_INVALID_VALUE, "grid order exceeds the supported maximum");
    if (mem != ZGL_MEM_DEVICE && mem != ZGL_MEM_HOST) return fail(ZGL_ERROR_INVALID_VALUE, "unknown memory kind");
    if (std::is_same_v<G, float> && plan->precision == ZGL_FP64)
        return fail(ZGL_ERROR_NOT_SUPPORTED, "FP32 grids need an FP32 plan");
    const size_t nt = std::min(exp_order, grid_order);
    const long gsz = (long)grid_size(plan, grid_order);
    const long csz = 2 * (long)coeff_size(plan, exp_order);
    const long gs = grid_stride ? (long)grid_stride : gsz;
    const long cs = coeff_stride ? (long)coeff_stride : csz;
    if (batch > 1 && gs < gsz) return fail(ZGL_ERROR_INVALID_VALUE, "instance strides overlap");
    return guarded([&]() -> zgl_status {
        if (grid_order) tables_for(plan, grid_order);
        plan->order = grid_order;
        if (batch == 0 || grid_order == 0) return ZGL_SUCCESS;
        if (!grid || (!coeffs && nt > 0)) return fail(ZGL_ERROR_INVALID_VALUE, "NULL grid or coefficient pointer");
        if (misaligned(grid, coeffs)) return fail(ZGL_ERROR_INVALID_VALUE, kMisaligned);
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (mem == ZGL_MEM_DEVICE) {
            const Inst<const double> ci{coeffs, cs, nullptr};
            const Inst<G> gi{grid, gs, nullptr};
            if constexpr (std::is_same_v<G, double>) {
                if (adjoint) forward_adjoint_any(plan, plan->main, ci, nt, gi, grid_order, batch, s);
                else backward_any<G>(plan, plan->main, ci, nt, gi, grid_order, batch, s);
            } else {
                backward_any<G>(plan, plan->main, ci, nt, gi, grid_order, batch, s);
            }
            check_hip(hipGetLastError(), "backward");
            return ZGL_SUCCESS;
        }
        const size_t ncr = 2 * coeff_size(plan, nt);  // doubles read per instance
        ensure(plan->dgrid, plan->dgrid_cap, ((size_t)gsz * batch * sizeof(G) + 7) / 8);
        ensure(plan->dcoef, plan->dcoef_cap, std::max<size_t>(ncr, 1) * batch);
        G* dg = reinterpret_cast<G*>(plan->dgrid);
        const Inst<const double> dci{plan->dcoef, (long)ncr, nullptr};
        const Inst<G> dgi{dg, gsz, nullptr};
        if (ncr)
            check_hip(hipMemcpy2DAsync(plan->dcoef, ncr * sizeof(double), coeffs, cs * sizeof(double),
                                       ncr * sizeof(double), batch, hipMemcpyHostToDevice, s),
                      "coefficient upload");
        if constexpr (std::is_same_v<G, double>) {
            if (adjoint) forward_adjoint_any(plan, plan->main, dci, nt, dgi, grid_order, batch, s);
            else backward_any<G>(plan, plan->main, dci, nt, dgi, grid_order, batch, s);
        } else {
            backward_any<G>(plan, plan->main, dci, nt, dgi, grid_order, batch, s);
        }
        check_hip(hipMemcpy2DAsync(grid, gs * sizeof(G), dg, gsz * sizeof(G), gsz * sizeof(G), batch,
                                   hipMemcpyDeviceToHost, s),
                  "grid download");
        check_hip(hipStreamSynchronize(s), "backward");
        return ZGL_SUCCESS;
    });
}

// ------------------------------------------------------------ ragged batches
// Jobs of uneven orders, two pipelines; both give the same bits as per-order
// batched calls.
//
// Lanes (the default): the jobs form an order-sorted queue of per-order
// groups (largest order first), dealt to kLanes internal streams by longest
// processing time on a modelled cost; each group runs the batched pipeline of
// its order with per-instance addresses (Inst offsets).
//
// Persistent (ZGL_RAGGED=persistent; Zernike, FP64 and FP32, where every
// order has the streaming kernels): the order-sorted queue is cut into waves
// whose spectra and flm fit the scratch budget. Each stage of a wave is one
// persistent launch over all the groups of the wave -- the FFT one per kernel
// variant among them -- whose work items (group, tile, instance group) are
// dealt to the resident workgroups by longest processing time on a modelled
// cost.
//
// Either way, on a non-null stream the whole pipeline is captured once into a
// hipGraph and replayed while the job list is unchanged (ZGL_RAGGED_GRAPH=0
// runs it without a graph); the persistent launch arrays are kept for the job
// list either way.
constexpr int kLanes = 4;

// The internal streams and their fork/join events, created on first use.
void ensure_lanes(zgl_plan::Ragged& R) {
    if (!R.lanes.empty()) return;
    R.lanes.resize(kLanes, nullptr);
    R.ev.resize(kLanes + 1, nullptr);
    R.scratch.resize(kLanes);
    for (auto& l : R.lanes) check_hip(hipStreamCreateWithFlags(&l, hipStreamNonBlocking), "lane stream");
    for (auto& e : R.ev) check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "lane event");
}

// Streams of the persistent pipeline: its waves alternate between them
// (ZGL_RAGGED_LANES, 1 to kLanes; default 1). A persistent launch fills every
// resident workgroup slot, so a second stream gains nothing unless launches
// are cut short (measured: 1 lane 102.9 ms, 2 lanes 105.0 ms per FP64 step of
// 10,000 instances of orders 8..96).
int persistent_lanes() {
    const char* e = std::getenv("ZGL_RAGGED_LANES");
    const int v = e && *e ? std::atoi(e) : 1;
    return std::max(1, std::min(kLanes, v));
}

// The pipeline ragged calls take: ZGL_RAGGED=persistent selects the
// persistent one, anything else the lanes one (measured faster: its per-order
// launches of different stages overlap on the four streams, which a launch
// filling the whole device cannot).
bool ragged_persistent_selected() {
    const char* e = std::getenv("ZGL_RAGGED");
    return e && std::strcmp(e, "persistent") == 0;
}

struct RaggedGroup {
    size_t order, first, count;
    int lane;
};

// Builds the persistent pipeline of one direction into C.plist: false (and
// nothing kept) where a group has no persistent form. `moved` reports a
// reallocated wave scratch (the other direction's pipeline points at the old).
bool build_persistent(zgl_plan* p, zgl_plan::Ragged::Cached& C, const std::vector<RaggedGroup>& groups,
                      double* gbase, double* cbase, size_t n, bool forward, bool& moved) {
    moved = false;
    if (p->kind != ZGL_ZERNIKE || (p->precision != ZGL_FP64 && p->precision != ZGL_FP32) ||
        !ragged_persistent_selected())
        return false;
    const bool fp32 = p->precision == ZGL_FP32;
    const size_t csize = fp32 ? sizeof(float2) : sizeof(double2);
    for (const auto& g : groups) {
        const OrderTables& t = tables_for(p, g.order);
        if (t.plm_in_kernel || (fp32 ? use_fused<float>(p, t) : use_fused<double>(p, t))) return false;
    }
    // waves: pieces of groups whose scratch fits the budget
    struct Piece {
        size_t order, first, count;
        size_t spec_off, flm_off;  // bytes into the wave scratch
    };
    std::vector<std::vector<Piece>> waves(1);
    const int nl = persistent_lanes();
    const size_t budget = (env_mb("ZGL_SCRATCH_MB", 4096) << 20) / nl;  // per lane
    size_t spec_use = 0, flm_use = 0, spec_max = 0, flm_max = 0;
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    for (const auto& g : groups) {
        const Chunking ch = chunking(p, tables_for(p, g.order).g, 1, csize);
        const size_t per_spec = (size_t)ch.spec_stride * csize, per_flm = (size_t)ch.flm_stride * csize;
        size_t done = 0;
        while (done < g.count) {
            const size_t room = budget > spec_use + flm_use ? budget - spec_use - flm_use : 0;
            size_t take = std::min(g.count - done, room / (per_spec + per_flm + 512));
            if (take == 0) {
                if (!waves.back().empty()) {  // start a new wave
                    waves.emplace_back();
                    spec_use = flm_use = 0;
                    continue;
                }
                take = 1;  // one instance above the budget: a wave of its own
            }
            waves.back().push_back(Piece{g.order, g.first + done, take, spec_use, flm_use});
            spec_use = align(spec_use + take * per_spec);
            flm_use = align(flm_use + take * per_flm);
            spec_max = std::max(spec_max, spec_use);
            flm_max = std::max(flm_max, flm_use);
            done += take;
        }
    }
    // wave w runs on lane w % nl, each lane on a scratch of its own
    auto& R = p->rg;
    if (nl > 1) ensure_lanes(R);
    if ((int)R.pscratch.size() < nl) R.pscratch.resize(nl);
    for (int l = 0; l < nl; ++l) {
        Scratch& sc = R.pscratch[l];
        const void *spec0 = sc.spec, *flm0 = sc.flm;
        ensure_bytes(sc.spec, sc.spec_cap, spec_max);
        ensure_bytes(sc.flm, sc.flm_cap, flm_max);
        moved = moved || sc.spec != spec0 || sc.flm != flm0;
    }

    std::vector<PersistentLaunch> out;
    try {
        for (size_t w = 0; w < waves.size(); ++w) {
            const auto& wave = waves[w];
            const int lane = (int)(w % nl);
            Scratch& sc = R.pscratch[lane];
            const size_t first_launch = out.size();
            std::vector<FftGroup> fg;
            std::vector<ZtGemmArgs> ga;
            for (const Piece& pc : wave) {
                const OrderTables& t = tables_for(p, pc.order);
                const Geom& g = t.g;
                const Chunking ch = chunking(p, g, 1, csize);
                char* spec = static_cast<char*>(sc.spec) + pc.spec_off;
                char* flm = static_cast<char*>(sc.flm) + pc.flm_off;
                FftLaunch L;
                L.g = &g;
                L.t = &t;
                L.layout_kind = GRID_COLS;
                L.grid_aligned16 = p->rg.grid16;
                L.nm_used = (int)pc.order;
                L.weighted = forward;
                L.scale = forward ? forward_prefactor(p, g) : 1.0;
                L.center_factor = 1.0;
                fg.push_back(FftGroup{L, Inst<double>{gbase, 0, C.offs + pc.first}, spec, ch.spec_stride,
                                      (int)pc.count});
                ZtGemmArgs a{&t, (int)pc.order, forward && p->znorm == ZGL_ZNORM_UNNORMED ? 1 : 0, spec,
                             ch.spec_stride, flm, ch.flm_stride, Inst<double>{cbase, 0, C.offs + n + pc.first},
                             (int)pc.count};
                ga.push_back(a);
            }
            bool ok;
            if (forward) {
                ok = fp32 ? fft_forward_persistent<float>(out, fg) : fft_forward_persistent<double>(out, fg);
                ok = ok && (fp32 ? zt_legendre_forward_persistent<float>(out, ga)
                                 : zt_legendre_forward_persistent<double>(out, ga));
                ok = ok && (fp32 ? zt_radial_forward_persistent<float>(out, ga)
                                 : zt_radial_forward_persistent<double>(out, ga));
            } else {
                ok = fp32 ? zt_radial_backward_persistent<float>(out, ga)
                          : zt_radial_backward_persistent<double>(out, ga);
                ok = ok && (fp32 ? zt_legendre_backward_persistent<float>(out, ga)
                                 : zt_legendre_backward_persistent<double>(out, ga));
                ok = ok && (fp32 ? fft_backward_persistent<float>(out, fg) : fft_backward_persistent<double>(out, fg));
            }
            if (!ok) {
                release(out);
                return false;
            }
            for (size_t k = first_launch; k < out.size(); ++k) out[k].lane = nl > 1 ? lane : -1;
        }
    } catch (...) {
        release(out);
        throw;
    }
    C.plist = std::move(out);
    return true;
}

Stage stage_of(const char* name) {
    for (int k = 0; k < ST_COUNT; ++k)
        if (std::strcmp(kStageNames[k], name) == 0) return (Stage)k;
    return ST_FFT_FWD;
}

// Launches on the caller's stream (lane -1) or forked over the lanes and
// joined back.
void run_persistent(zgl_plan* p, const zgl_plan::Ragged::Cached& C, hipStream_t s) {
    auto& R = p->rg;
    bool forked = false;
    for (const PersistentLaunch& pl : C.plist) forked = forked || pl.lane >= 0;
    if (forked) {
        check_hip(hipEventRecord(R.ev[0], s), "fork");
        for (auto l : R.lanes) check_hip(hipStreamWaitEvent(l, R.ev[0], 0), "fork");
    }
    for (const PersistentLaunch& pl : C.plist) {
        hipStream_t ls = pl.lane >= 0 ? R.lanes[pl.lane] : s;
        if (p->profiling) timed(p, stage_of(pl.stage), (int)pl.instances, ls, [&] { pl.run(ls); });
        else pl.run(ls);
    }
    if (forked)
        for (int l = 0; l < kLanes; ++l) {
            check_hip(hipEventRecord(R.ev[1 + l], R.lanes[l]), "join");
            check_hip(hipStreamWaitEvent(s, R.ev[1 + l], 0), "join");
        }
}

bool ragged_graphs() {
    const char* e = std::getenv("ZGL_RAGGED_GRAPH");
    return !(e && e[0] == '0');
}

// modelled work of `count` instances of order N: memory traffic ~ N^3
double job_cost(size_t N, size_t count) {
    const double n = (double)N + 1.0;
    return (double)count * n * n * n;
}

// Waits for the last launch that used the cache (its offsets and graph).
void wait_cache(zgl_plan::Ragged::Cached& c) {
    if (c.done) check_hip(hipEventSynchronize(c.done), "ragged: wait for the previous call");
}

// Marks the end of a launch that uses the cache on stream s.
void mark_cache(zgl_plan::Ragged::Cached& c, hipStream_t s) {
    if (!c.done) check_hip(hipEventCreateWithFlags(&c.done, hipEventDisableTiming), "ragged event");
    check_hip(hipEventRecord(c.done, s), "ragged event");
}

void drop_graph(zgl_plan::Ragged::Cached& c) {
    wait_cache(c);
    if (c.exec) (void)hipGraphExecDestroy(c.exec);
    if (c.graph) (void)hipGraphDestroy(c.graph);
    c.exec = nullptr;
    c.graph = nullptr;
    c.jobs.clear();
    c.valid = false;
    release(c.plist);
    c.persistent = false;
}

bool same_jobs(const std::vector<zgl_job>& a, const zgl_job* b, size_t n) {
    if (a.size() != n) return false;
    for (size_t i = 0; i < n; ++i)
        if (a[i].order != b[i].order || a[i].grid != b[i].grid || a[i].coeffs != b[i].coeffs) return false;
    return true;
}

// scratch for `count` instances of order N in the plan's precision; true
// when a buffer was (re)allocated
bool reserve(zgl_plan* p, Scratch& sc, size_t N, size_t count) {
    const Geom& g = tables_for(p, N).g;
    const Scratch before = sc;
    for (size_t csize : {sizeof(double2), sizeof(float2)}) {
        if (csize == sizeof(double2) && (p->precision == ZGL_FP32 || p->precision == ZGL_BF16X3)) continue;
        if (csize == sizeof(float2) && p->precision == ZGL_FP64) continue;
        const Chunking ch = chunking(p, g, count, csize);
        ensure_bytes(sc.spec, sc.spec_cap, (size_t)ch.spec_stride * ch.nb * csize);
        if (ch.flm_stride) ensure_bytes(sc.flm, sc.flm_cap, (size_t)ch.flm_stride * ch.nb * csize);
    }
    if (p->precision == ZGL_FP32_REFINE) {
        const Chunking ch = chunking(p, g, count, sizeof(float2));
        ensure(sc.rgrid, sc.rgrid_cap, grid_size(p, N) * ch.nb);
        ensure(sc.rcoef, sc.rcoef_cap, 2 * coeff_size(p, N) * ch.nb);
    }
    return sc.spec != before.spec || sc.flm != before.flm || sc.rgrid != before.rgrid ||
           sc.rcoef != before.rcoef;
}

zgl_status ragged_entry(zgl_plan* plan, const zgl_job* jobs, size_t njobs, void* stream, bool forward) {
    if (!plan) return fail(ZGL_ERROR_INVALID_VALUE, "plan is NULL");
    if (njobs && !jobs) return fail(ZGL_ERROR_INVALID_VALUE, "jobs is NULL");
    for (size_t i = 0; i < njobs; ++i) {
        if (jobs[i].order > kMaxOrder) return fail(ZGL_ERROR_INVALID_VALUE, "job order exceeds the supported maximum");
        if (jobs[i].order && (!jobs[i].grid || !jobs[i].coeffs))
            return fail(ZGL_ERROR_INVALID_VALUE, "NULL grid or coefficient pointer in a job");
        if (jobs[i].order && misaligned(static_cast<const double*>(jobs[i].grid),
                                        static_cast<const double*>(jobs[i].coeffs)))
            return fail(ZGL_ERROR_INVALID_VALUE, std::string(kMisaligned) + " in a job");
    }
    return guarded([&]() -> zgl_status {
        hipStream_t s = static_cast<hipStream_t>(stream);
        auto& R = plan->rg;
        auto& C = R.cache[forward ? 1 : 0];
        const bool use_graph = s != nullptr && !plan->profiling && ragged_graphs();
        const bool same = C.valid && same_jobs(C.jobs, jobs, njobs);
        if (same && use_graph && C.exec) {
            check_hip(hipGraphLaunch(C.exec, s), "ragged graph launch");
            mark_cache(C, s);
            return ZGL_SUCCESS;
        }
        if (same && !use_graph && C.persistent) {
            run_persistent(plan, C, s);
            mark_cache(C, s);
            return ZGL_SUCCESS;
        }
        // the previous call of this direction may still be reading C.offs or
        // running C.exec: drop_graph waits for it before either changes
        drop_graph(C);

        // order-sorted queue of non-empty jobs, largest order first
        std::vector<size_t> idx;
        for (size_t i = 0; i < njobs; ++i)
            if (jobs[i].order) idx.push_back(i);
        if (idx.empty()) return ZGL_SUCCESS;
        std::stable_sort(idx.begin(), idx.end(),
                         [&](size_t a, size_t b) { return jobs[a].order > jobs[b].order; });
        std::vector<RaggedGroup> groups;
        for (size_t k = 0; k < idx.size(); ++k) {
            if (groups.empty() || groups.back().order != jobs[idx[k]].order)
                groups.push_back(RaggedGroup{jobs[idx[k]].order, k, 0, 0});
            ++groups.back().count;
        }

        // per-instance element offsets from the lowest grid / coefficient address
        const size_t n = idx.size();
        R.grid16 = true;
        for (size_t k : idx) R.grid16 = R.grid16 && reinterpret_cast<uintptr_t>(jobs[k].grid) % 16 == 0;
        double* gbase = jobs[idx[0]].grid;
        double* cbase = jobs[idx[0]].coeffs;
        for (size_t k : idx) {
            gbase = std::min(gbase, jobs[k].grid);
            cbase = std::min(cbase, jobs[k].coeffs);
        }
        std::vector<long> offs(2 * n);
        for (size_t k = 0; k < n; ++k) {
            offs[k] = (long)(jobs[idx[k]].grid - gbase);
            offs[n + k] = (long)(jobs[idx[k]].coeffs - cbase);
        }
        ensure(C.offs, C.offs_cap, 2 * n);
        check_hip(hipMemcpy(C.offs, offs.data(), 2 * n * sizeof(long), hipMemcpyHostToDevice),
                  "ragged offsets upload");

        bool moved = false;
        if (build_persistent(plan, C, groups, gbase, cbase, n, forward, moved)) {
            // the other direction's pipeline points at the old wave scratch
            if (moved) drop_graph(R.cache[forward ? 0 : 1]);
            C.persistent = true;
            C.jobs.assign(jobs, jobs + njobs);
            C.valid = true;
            R.lane_cost.clear();
            double best = -1;
            for (const auto& pl : C.plist) {  // the launch of the most modelled work
                const double tot = pl.cost_mean * pl.bin_cost.size();
                if (tot > best) {
                    best = tot;
                    R.lane_cost = pl.bin_cost;
                }
            }
            if (!use_graph) {
                run_persistent(plan, C, s);
                mark_cache(C, s);
                return ZGL_SUCCESS;
            }
            check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), "begin capture");
            try {
                run_persistent(plan, C, s);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            check_hip(hipStreamEndCapture(s, &C.graph), "end capture");
            check_hip(hipGraphInstantiate(&C.exec, C.graph, nullptr, nullptr, 0), "graph instantiate");
            check_hip(hipGraphLaunch(C.exec, s), "ragged graph launch");
            mark_cache(C, s);
            return ZGL_SUCCESS;
        }

        // lanes: streams, events, scratch; groups dealt by longest processing time
        ensure_lanes(R);
        R.lane_cost.assign(kLanes, 0.0);
        moved = false;
        for (auto& g : groups) {
            tables_for(plan, g.order);
            const int l = (int)(std::min_element(R.lane_cost.begin(), R.lane_cost.end()) - R.lane_cost.begin());
            g.lane = l;
            R.lane_cost[l] += job_cost(g.order, g.count);
            moved |= reserve(plan, R.scratch[l], g.order, g.count);
        }
        // the other direction's graph points at the old scratch
        if (moved) drop_graph(R.cache[forward ? 0 : 1]);

        auto enqueue = [&]() {
            check_hip(hipEventRecord(R.ev[0], s), "fork");
            for (auto l : R.lanes) check_hip(hipStreamWaitEvent(l, R.ev[0], 0), "fork");
            for (const auto& g : groups) {
                hipStream_t ls = R.lanes[g.lane];
                Scratch& sc = R.scratch[g.lane];
                const long* go = C.offs + g.first;
                const long* co = C.offs + n + g.first;
                if (forward)
                    forward_any<double>(plan, sc, Inst<const double>{gbase, 0, go}, g.order,
                                        Inst<double>{cbase, 0, co}, g.order, g.count, ls);
                else
                    backward_any<double>(plan, sc, Inst<const double>{cbase, 0, co}, g.order,
                                         Inst<double>{gbase, 0, go}, g.order, g.count, ls);
            }
            for (int l = 0; l < kLanes; ++l) {
                check_hip(hipEventRecord(R.ev[1 + l], R.lanes[l]), "join");
                check_hip(hipStreamWaitEvent(s, R.ev[1 + l], 0), "join");
            }
        };

        if (!use_graph) {
            enqueue();
            check_hip(hipGetLastError(), "ragged");
            mark_cache(C, s);
            return ZGL_SUCCESS;
        }
        check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), "begin capture");
        try {
            enqueue();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        check_hip(hipStreamEndCapture(s, &C.graph), "end capture");
        check_hip(hipGraphInstantiate(&C.exec, C.graph, nullptr, nullptr, 0), "graph instantiate");
        C.jobs.assign(jobs, jobs + njobs);
        C.valid = true;
        check_hip(hipGraphLaunch(C.exec, s), "ragged graph launch");
        mark_cache(C, s);
        return ZGL_SUCCESS;
    });
}

}  // namespace
}  // namespace zgl

extern "C" {

zgl_status zgl_forward(zgl_plan* plan, const double* grid, size_t grid_order, double* coeffs,
                       size_t exp_order, size_t batch, ptrdiff_t grid_stride,
                       ptrdiff_t coeff_stride, zgl_mem mem, void* stream) {
    return forward_entry<double>(plan, grid, grid_order, coeffs, exp_order, batch, grid_stride,
                                 coeff_stride, mem, stream);
}

zgl_status zgl_backward(zgl_plan* plan, const double* coeffs, size_t exp_order, double* grid,
                        size_t grid_order, size_t batch, ptrdiff_t coeff_stride,
                        ptrdiff_t grid_stride, zgl_mem mem, void* stream) {
    return backward_entry<double>(plan, coeffs, exp_order, grid, grid_order, batch, coeff_stride,
                                  grid_stride, mem, stream);
}

}  // extern "C"

// ------------------------------------------------- forward without allocation
// Device scratch of zgl_forward_noalloc: the spectrum, flm / staging and
// refinement buffers of one chunk of instances, carved out of this array.
// Every GPU the library runs on holds its own copy until the process ends.
// In host code `scratch` is only a stand-in for the device address: the
// address on the current device comes from hipGetSymbolAddress, resolved once
// per device (at plan creation, or on the entry point's first call there).
static constexpr size_t kStaticScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace zgl {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device.
void* static_scratch() {
    int dev = 0;
    check_hip(hipGetDevice(&dev), "current device");
    if (dev < 0 || dev >= kMaxDevices)
        throw Error{ZGL_ERROR_NOT_SUPPORTED, "device index beyond the static scratch table"};
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "static scratch address");
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

void resolve_static_scratch() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return;  // no device: resolved on the entry point's first call instead
    }
    try {
        (void)static_scratch();
    } catch (const Error&) {
        (void)hipGetLastError();  // reported by the entry point, if it is used
    }
}

// For the duration of a no-allocation call: no autotuning or stage timing
// (both create events), chunks capped to what the static scratch holds, and
// any buffer growth or table build turned into an error.
struct NoAllocScope {
    zgl_plan* p;
    bool autotune, profiling;
    NoAllocScope(zgl_plan* plan, size_t cap) : p(plan), autotune(plan->autotune), profiling(plan->profiling) {
        p->autotune = false;
        p->profiling = false;
        p->chunk_cap = cap;
        t_no_alloc = true;
    }
    ~NoAllocScope() {
        t_no_alloc = false;
        p->chunk_cap = 0;
        p->autotune = autotune;
        p->profiling = profiling;
    }
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace
}  // namespace zgl

extern "C" zgl_status zgl_forward_noalloc(zgl_plan* plan, const double* grid, size_t grid_order,
                                          double* coeffs, size_t exp_order, size_t batch,
                                          ptrdiff_t grid_stride, ptrdiff_t coeff_stride, void* stream) {
    using namespace zgl;
    if (!plan) return fail(ZGL_ERROR_INVALID_VALUE, "plan is NULL");
    if (grid_order > kMaxOrder) return fail(ZGL_ERROR_INVALID_VALUE, "grid order exceeds the supported maximum");
    const size_t nt = std::min(exp_order, grid_order);
    const long gsz = (long)grid_size(plan, grid_order);
    const long csz = 2 * (long)coeff_size(plan, exp_order);
    const long gs = grid_stride ? (long)grid_stride : gsz;
    const long cs = coeff_stride ? (long)coeff_stride : csz;
    if (batch > 1 && (gs < gsz || cs < 2 * (long)coeff_size(plan, nt)))
        return fail(ZGL_ERROR_INVALID_VALUE, "instance strides overlap");
    if (grid_order && plan->tables.find(grid_order) == plan->tables.end())
        return fail(ZGL_ERROR_INVALID_VALUE,
                    "no tables for this grid order: create or resize the plan to it before the "
                    "no-allocation call");
    return guarded([&]() -> zgl_status {
        plan->order = grid_order;
        if (batch == 0 || nt == 0) return ZGL_SUCCESS;
        if (!grid || !coeffs) return fail(ZGL_ERROR_INVALID_VALUE, "NULL grid or coefficient pointer");
        if (misaligned(grid, coeffs)) return fail(ZGL_ERROR_INVALID_VALUE, kMisaligned);
        const Geom& g = plan->tables.find(grid_order)->second.g;
        // bytes of scratch per instance: spectrum and flm (SH: staging) in the
        // widest arithmetic the call runs, and the refinement buffers
        const bool wide = plan->precision == ZGL_FP64 || plan->precision == ZGL_FP32_REFINE;
        const size_t csize = wide ? sizeof(double2) : sizeof(float2);
        const Chunking one = chunking(plan, g, 1, csize);
        const size_t spec_b = (size_t)one.spec_stride * csize, flm_b = (size_t)one.flm_stride * csize;
        const size_t ncw = 2 * coeff_size(plan, nt);
        const bool refine = plan->precision == ZGL_FP32_REFINE;
        const size_t per = spec_b + flm_b + (refine ? sizeof(double) * ((size_t)gsz + ncw) : 0);
        const size_t nb_max = std::min<size_t>(batch, (kStaticScratchBytes - 4 * 256) / per);
        if (nb_max == 0)
            return fail(ZGL_ERROR_INVALID_VALUE,
                        "one instance of this order needs " + std::to_string(per) +
                            " bytes of scratch; the no-allocation path holds " +
                            std::to_string(kStaticScratchBytes) + " (zgl_forward has no such limit)");
        unsigned char* base = static_cast<unsigned char*>(static_scratch());
        Scratch sc;
        size_t off = 0;
        sc.spec = base + off; sc.spec_cap = nb_max * spec_b; off = align256(off + sc.spec_cap);
        sc.flm = base + off; sc.flm_cap = nb_max * flm_b; off = align256(off + sc.flm_cap);
        if (refine) {
            sc.rgrid = reinterpret_cast<double*>(base + off); sc.rgrid_cap = nb_max * (size_t)gsz;
            off = align256(off + sc.rgrid_cap * sizeof(double));
            sc.rcoef = reinterpret_cast<double*>(base + off); sc.rcoef_cap = nb_max * ncw;
            off = align256(off + sc.rcoef_cap * sizeof(double));
        }
        if (off > kStaticScratchBytes) throw Error{ZGL_ERROR_HIP, "static scratch layout (internal)"};
        hipStream_t s = static_cast<hipStream_t>(stream);
        {
            NoAllocScope scope(plan, nb_max);
            forward_any<double>(plan, sc, Inst<const double>{grid, gs, nullptr}, grid_order,
                                Inst<double>{coeffs, cs, nullptr}, nt, batch, s);
        }
        check_hip(hipGetLastError(), "forward (no allocation)");
        return ZGL_SUCCESS;
    });
}

extern "C" {

zgl_status zgl_forward_ragged(zgl_plan* plan, const zgl_job* jobs, size_t njobs, void* stream) {
    return ragged_entry(plan, jobs, njobs, stream, true);
}

zgl_status zgl_backward_ragged(zgl_plan* plan, const zgl_job* jobs, size_t njobs, void* stream) {
    return ragged_entry(plan, jobs, njobs, stream, false);
}

zgl_status zgl_ragged_lane_costs(const zgl_plan* plan, double* costs, int max_lanes, int* num_lanes) {
    if (!plan || !num_lanes) return fail(ZGL_ERROR_INVALID_VALUE, "NULL argument");
    const auto& c = plan->rg.lane_cost;
    *num_lanes = (int)c.size();
    for (int i = 0; i < (int)c.size() && i < max_lanes && costs; ++i) costs[i] = c[i];
    return ZGL_SUCCESS;
}

zgl_status zgl_backward_adjoint(zgl_plan* plan, const double* grid, size_t grid_order,
                                double* coeffs, size_t exp_order, size_t batch,
                                ptrdiff_t grid_stride, ptrdiff_t coeff_stride, zgl_mem mem,
                                void* stream) {
    return forward_entry<double>(plan, grid, grid_order, coeffs, exp_order, batch, grid_stride,
                                 coeff_stride, mem, stream, true);
}

zgl_status zgl_forward_adjoint(zgl_plan* plan, const double* coeffs, size_t exp_order,
                               double* grid, size_t grid_order, size_t batch,
                               ptrdiff_t coeff_stride, ptrdiff_t grid_stride, zgl_mem mem,
                               void* stream) {
    return backward_entry<double>(plan, coeffs, exp_order, grid, grid_order, batch, coeff_stride,
                                  grid_stride, mem, stream, true);
}

zgl_status zgl_forward_grid32(zgl_plan* plan, const float* grid, size_t grid_order, double* coeffs,
                              size_t exp_order, size_t batch, ptrdiff_t grid_stride,
                              ptrdiff_t coeff_stride, zgl_mem mem, void* stream) {
    return forward_entry<float>(plan, grid, grid_order, coeffs, exp_order, batch, grid_stride,
                                coeff_stride, mem, stream);
}

zgl_status zgl_backward_grid32(zgl_plan* plan, const double* coeffs, size_t exp_order, float* grid,
                               size_t grid_order, size_t batch, ptrdiff_t coeff_stride,
                               ptrdiff_t grid_stride, zgl_mem mem, void* stream) {
    return backward_entry<float>(plan, coeffs, exp_order, grid, grid_order, batch, coeff_stride,
                                 grid_stride, mem, stream);
}

zgl_status zgl_plan_autotune(zgl_plan* plan, int enable) {
    if (!plan) return fail(ZGL_ERROR_INVALID_VALUE, "plan is NULL");
    plan->autotune = enable != 0;
    return ZGL_SUCCESS;
}

zgl_status zgl_plan_tuning(const zgl_plan* plan, int* entries, int max_entries, int* num_entries) {
    if (!plan || !num_entries) return fail(ZGL_ERROR_INVALID_VALUE, "NULL argument");
    int i = 0;
    for (const auto& kv : plan->tuned) {
        if (entries && i < max_entries) {
            entries[5 * i + 0] = (int)std::get<0>(kv.first);
            entries[5 * i + 1] = std::get<1>(kv.first);
            entries[5 * i + 2] = std::get<3>(kv.first);
            entries[5 * i + 3] = std::get<2>(kv.first);
            entries[5 * i + 4] = kv.second;
        }
        ++i;
    }
    *num_entries = i;
    return ZGL_SUCCESS;
}

zgl_status zgl_profile_enable(zgl_plan* plan, int enable) {
    if (!plan) return fail(ZGL_ERROR_INVALID_VALUE, "plan is NULL");
    plan->profiling = enable != 0;
    return ZGL_SUCCESS;
}

zgl_status zgl_profile_stats(zgl_plan* plan, zgl_stage_stats* out, int max_stages, int* num_stages) {
    if (!plan || !num_stages || (max_stages > 0 && !out))
        return fail(ZGL_ERROR_INVALID_VALUE, "NULL argument");
    return guarded([&]() -> zgl_status {
        long long launches[ST_COUNT] = {}, inst[ST_COUNT] = {};
        double ms[ST_COUNT] = {};
        for (auto& r : plan->recs) {
            check_hip(hipEventSynchronize(r.stop), "event wait");
            float t = 0.f;
            check_hip(hipEventElapsedTime(&t, r.start, r.stop), "event time");
            launches[r.stage] += 1;
            inst[r.stage] += r.nb;
            ms[r.stage] += t;
            plan->pool.push_back(r.start);
            plan->pool.push_back(r.stop);
        }
        plan->recs.clear();
        int n = 0;
        for (int st = 0; st < ST_COUNT; ++st) {
            if (!launches[st]) continue;
            if (n < max_stages) {
                zgl_stage_stats& o = out[n];
                std::snprintf(o.name, sizeof(o.name), "%s", kStageNames[st]);
                o.launches = launches[st];
                o.instances = inst[st];
                o.total_ms = ms[st];
            }
            ++n;
        }
        *num_stages = n;
        return ZGL_SUCCESS;
    });
}

const char* zgl_last_error_string(void) { return g_last_error.c_str(); }

zgl_status zgl_set_device_allocator(zgl_device_alloc_fn alloc, zgl_device_free_fn free, void* user) {
    if (alloc && !free) return fail(ZGL_ERROR_INVALID_VALUE, "an allocator needs its free function");
    std::lock_guard<std::mutex> lock(g_alloc_mutex);
    g_allocator = alloc ? DeviceAllocator{alloc, free, user} : DeviceAllocator{};
    return ZGL_SUCCESS;
}

}  // extern "C"

// ------------------------------------------------------------- evaluation
extern "C" zgl_status zgl_evaluate(zgl_plan* plan, const double* coeffs, size_t exp_order,
                                   const double* longitudes, size_t nlon,
                                   const double* colatitudes, size_t nlat, const double* radii,
                                   size_t nrad, double* values, size_t batch,
                                   ptrdiff_t coeff_stride, ptrdiff_t value_stride, zgl_mem mem,
                                   void* stream) {
    using namespace zgl;
    if (!plan) return fail(ZGL_ERROR_INVALID_VALUE, "plan is NULL");
    if (exp_order > kMaxOrder) return fail(ZGL_ERROR_INVALID_VALUE, "order exceeds the supported maximum");
    if (mem != ZGL_MEM_DEVICE && mem != ZGL_MEM_HOST) return fail(ZGL_ERROR_INVALID_VALUE, "unknown memory kind");
    const bool zt = plan->kind == ZGL_ZERNIKE;
    if (!zt) nrad = 1;
    const size_t kMaxPoints = (size_t)1 << 30;
    if (nlon > kMaxPoints || nlat > kMaxPoints || nrad > kMaxPoints || batch > kMaxPoints)
        return fail(ZGL_ERROR_INVALID_VALUE, "point or instance count too large");
    // the point count must fit the kernels' signed 64-bit indexing with room
    // for the instance offsets: at most 2^40 points per instance (each factor
    // is below 2^30, so the checks themselves cannot overflow)
    const size_t kMaxGrid = (size_t)1 << 40;
    if (nlon && nlat && (nlat > kMaxGrid / nlon || (nrad && nrad > kMaxGrid / (nlon * nlat))))
        return fail(ZGL_ERROR_INVALID_VALUE, "evaluation grid has too many points");
    const long npts = (long)(nlon * nlat * nrad);
    const long csz = 2 * (long)coeff_size(plan, exp_order);
    const long cs = coeff_stride ? (long)coeff_stride : csz;
    const long vs = value_stride ? (long)value_stride : npts;
    if (batch > 1 && (vs < npts || cs < csz)) return fail(ZGL_ERROR_INVALID_VALUE, "instance strides overlap");
    if (npts == 0 || batch == 0) return ZGL_SUCCESS;
    if (!values || !longitudes || !colatitudes || (zt && !radii) || (!coeffs && exp_order > 0))
        return fail(ZGL_ERROR_INVALID_VALUE, "NULL pointer");
    if (misaligned(values, longitudes, colatitudes, radii, coeffs)) return fail(ZGL_ERROR_INVALID_VALUE, kMisaligned);
    return guarded([&]() -> zgl_status {
        hipStream_t s = static_cast<hipStream_t>(stream);
        EvalCall e{plan->kind, plan->znorm, plan->shnorm == ZGL_SHNORM_QM ? 1 : 0, (double)plan->phase,
                   (int)exp_order, Inst<const double>{coeffs, cs, nullptr},
                   longitudes, (int)nlon, colatitudes, (int)nlat, zt ? radii : nullptr, (int)nrad,
                   values, vs, (int)batch};
        EvalTimer timer = nullptr;
        if (plan->profiling)
            timer = [&](int stage, int nb, const std::function<void()>& f) {
                timed(plan, (Stage)(ST_EVAL_TABLES + stage), nb, s, f);
            };
        if (mem == ZGL_MEM_DEVICE) {
            evaluate(plan->eval, e, s, timer);
            check_hip(hipGetLastError(), "evaluate");
            return ZGL_SUCCESS;
        }
        const size_t np = nlon + nlat + (zt ? nrad : 0);
        ensure(plan->dpts, plan->dpts_cap, np);
        ensure(plan->dcoef, plan->dcoef_cap, std::max<size_t>((size_t)csz, 1) * batch);
        ensure(plan->dgrid, plan->dgrid_cap, (size_t)npts * batch);
        double* dp = plan->dpts;
        check_hip(hipMemcpyAsync(dp, longitudes, nlon * sizeof(double), hipMemcpyHostToDevice, s), "points");
        check_hip(hipMemcpyAsync(dp + nlon, colatitudes, nlat * sizeof(double), hipMemcpyHostToDevice, s),
                  "points");
        if (zt)
            check_hip(hipMemcpyAsync(dp + nlon + nlat, radii, nrad * sizeof(double), hipMemcpyHostToDevice, s),
                      "points");
        if (csz)
            check_hip(hipMemcpy2DAsync(plan->dcoef, csz * sizeof(double), coeffs, cs * sizeof(double),
                                       csz * sizeof(double), batch, hipMemcpyHostToDevice, s),
                      "coefficient upload");
        e.coeffs = Inst<const double>{plan->dcoef, csz, nullptr};
        e.lon = dp;
        e.colat = dp + nlon;
        e.radii = zt ? dp + nlon + nlat : nullptr;
        e.values = plan->dgrid;
        e.value_stride = npts;
        evaluate(plan->eval, e, s, timer);
        check_hip(hipMemcpy2DAsync(values, vs * sizeof(double), plan->dgrid, npts * sizeof(double),
                                   npts * sizeof(double), batch, hipMemcpyDeviceToHost, s),
                  "value download");
        check_hip(hipStreamSynchronize(s), "evaluate");
        return ZGL_SUCCESS;
    });
}

// ------------------------------------------------------------ power spectra
namespace zgl {
namespace {
zgl_status spectrum_entry(zgl_plan* plan, const double* a, const double* b, size_t order, double* out,
                          size_t batch, ptrdiff_t coeff_stride, ptrdiff_t out_stride, zgl_mem mem,
                          void* stream) {
    if (!plan) return fail(ZGL_ERROR_INVALID_VALUE, "plan is NULL");
    if (order > kMaxOrder) return fail(ZGL_ERROR_INVALID_VALUE, "order exceeds the supported maximum");
    if (mem != ZGL_MEM_DEVICE && mem != ZGL_MEM_HOST) return fail(ZGL_ERROR_INVALID_VALUE, "unknown memory kind");
    const long csz = 2 * (long)coeff_size(plan, order);
    const long nent = spectrum_size(plan->kind, (int)order);
    const long cs = coeff_stride ? (long)coeff_stride : csz;
    const long os = out_stride ? (long)out_stride : nent;
    if (batch > 1 && (cs < csz || os < nent)) return fail(ZGL_ERROR_INVALID_VALUE, "instance strides overlap");
    if (batch > 65535) return fail(ZGL_ERROR_INVALID_VALUE, "batch exceeds 65535 instances per call");
    if (order == 0 || batch == 0) return ZGL_SUCCESS;
    if (!a || !b || !out) return fail(ZGL_ERROR_INVALID_VALUE, "NULL pointer");
    if (misaligned(a, b, out)) return fail(ZGL_ERROR_INVALID_VALUE, kMisaligned);
    return guarded([&]() -> zgl_status {
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (mem == ZGL_MEM_DEVICE) {
            timed(plan, ST_SPECTRUM, (int)batch, s, [&] {
                power_spectrum(plan->kind, plan->shnorm, (int)order, Inst<const double>{a, cs, nullptr},
                               Inst<const double>{b, cs, nullptr}, out, os, (int)batch, s);
            });
            return ZGL_SUCCESS;
        }
        const bool cross = a != b;
        ensure(plan->dcoef, plan->dcoef_cap, (size_t)csz * batch * (cross ? 2 : 1));
        ensure(plan->dgrid, plan->dgrid_cap, (size_t)nent * batch);
        double* da = plan->dcoef;
        double* db = cross ? plan->dcoef + (size_t)csz * batch : da;
        check_hip(hipMemcpy2DAsync(da, csz * sizeof(double), a, cs * sizeof(double), csz * sizeof(double),
                                   batch, hipMemcpyHostToDevice, s),
                  "coefficient upload");
        if (cross)
            check_hip(hipMemcpy2DAsync(db, csz * sizeof(double), b, cs * sizeof(double), csz * sizeof(double),
                                       batch, hipMemcpyHostToDevice, s),
                      "coefficient upload");
        power_spectrum(plan->kind, plan->shnorm, (int)order, Inst<const double>{da, csz, nullptr},
                       Inst<const double>{db, csz, nullptr}, plan->dgrid, nent, (int)batch, s);
        check_hip(hipMemcpy2DAsync(out, os * sizeof(double), plan->dgrid, nent * sizeof(double),
                                   nent * sizeof(double), batch, hipMemcpyDeviceToHost, s),
                  "spectrum download");
        check_hip(hipStreamSynchronize(s), "power spectrum");
        return ZGL_SUCCESS;
    });
}
}  // namespace
}  // namespace zgl

extern "C" zgl_status zgl_power_spectrum(zgl_plan* plan, const double* coeffs, size_t order, double* out,
                                         size_t batch, ptrdiff_t coeff_stride, ptrdiff_t out_stride,
                                         zgl_mem mem, void* stream) {
    return zgl::spectrum_entry(plan, coeffs, coeffs, order, out, batch, coeff_stride, out_stride, mem,
                               stream);
}

extern "C" zgl_status zgl_cross_power_spectrum(zgl_plan* plan, const double* a, const double* b,
                                               size_t order, double* out, size_t batch,
                                               ptrdiff_t coeff_stride, ptrdiff_t out_stride, zgl_mem mem,
                                               void* stream) {
    return zgl::spectrum_entry(plan, a, b, order, out, batch, coeff_stride, out_stride, mem, stream);
}

// ------------------------------------------------------------ rotations
namespace zgl {
namespace {
// The plan's d(pi/2) table covering degrees < order (grown, never shrunk).
const double* wigner_for(zgl_plan* p, int order, hipStream_t s) {
    if (p->wtab_order < order) {
        if (p->wtab) {
            check_hip(hipDeviceSynchronize(), "d table growth");
            dfree(p->wtab);
            p->wtab = nullptr;
            p->wtab_order = 0;
        }
        p->wtab = dmalloc_n<double>((size_t)std::max<long>(wignerd_size(order), 1), "d table allocation");
        wignerd_table(order, p->wtab, s);
        p->wtab_order = order;
    }
    return p->wtab;
}

std::pair<int4*, int> rot_work_for(zgl_plan* p, bool fused, int order, int nb, hipStream_t s) {
    const auto key = std::make_tuple(fused ? 1 : 0, (int)p->kind, order, nb);
    auto it = p->rot_work.find(key);
    if (it != p->rot_work.end()) return it->second;
    if (p->rot_work.size() >= 64) {
        // many distinct batch sizes: start the cache over
        check_hip(hipDeviceSynchronize(), "work list cache");
        for (auto& kv : p->rot_work) dfree(kv.second.first);
        p->rot_work.clear();
    }
    const std::vector<int4> w = fused ? rotation_fused_work(p->kind == ZGL_ZERNIKE, order, nb)
                                      : rotation_work(p->kind == ZGL_ZERNIKE, order, nb);
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
    return p->rot_work.emplace(key, std::make_pair(d, (int)w.size())).first->second;
}

zgl_status rotate_entry(zgl_plan* plan, double* coeffs, size_t order, const double* angles, int rtype,
                        size_t batch, ptrdiff_t coeff_stride, zgl_mem mem, void* stream, bool polar) {
    if (!plan) return fail(ZGL_ERROR_INVALID_VALUE, "plan is NULL");
    if (order > kMaxOrder) return fail(ZGL_ERROR_INVALID_VALUE, "order exceeds the supported maximum");
    if (mem != ZGL_MEM_DEVICE && mem != ZGL_MEM_HOST) return fail(ZGL_ERROR_INVALID_VALUE, "unknown memory kind");
    if (rtype != 0 && rtype != 1) return fail(ZGL_ERROR_INVALID_VALUE, "rotation type must be 0 (object) or 1 (coordinate)");
    const long csz = 2 * (long)coeff_size(plan, order);
    const long cs = coeff_stride ? (long)coeff_stride : csz;
    if (batch > 1 && cs < csz) return fail(ZGL_ERROR_INVALID_VALUE, "instance strides overlap");
    if (batch > 65535) return fail(ZGL_ERROR_INVALID_VALUE, "batch exceeds 65535 instances per call");
    if (order < 2 || batch == 0) return ZGL_SUCCESS;
    if (!coeffs || !angles) return fail(ZGL_ERROR_INVALID_VALUE, "NULL pointer");
    if (misaligned(coeffs, angles)) return fail(ZGL_ERROR_INVALID_VALUE, kMisaligned);
    const bool zt = plan->kind == ZGL_ZERNIKE;
    const size_t nang = (polar ? 1 : 3) * batch;
    return guarded([&]() -> zgl_status {
        hipStream_t s = static_cast<hipStream_t>(stream);
        double* dc = coeffs;
        long dcs = cs;
        const double* da = angles;
        if (mem == ZGL_MEM_HOST) {
            ensure(plan->dcoef, plan->dcoef_cap, (size_t)csz * batch);
            ensure(plan->dang, plan->dang_cap, nang);
            check_hip(hipMemcpy2DAsync(plan->dcoef, csz * sizeof(double), coeffs, cs * sizeof(double),
                                       csz * sizeof(double), batch, hipMemcpyHostToDevice, s),
                      "coefficient upload");
            check_hip(hipMemcpyAsync(plan->dang, angles, nang * sizeof(double), hipMemcpyHostToDevice, s),
                      "angle upload");
            dc = plan->dcoef;
            dcs = csz;
            da = plan->dang;
        }
        const Inst<double> ci{dc, dcs, nullptr};
        if (polar) {
            timed(plan, ST_POLAR, (int)batch, s, [&] { polar_rotate(zt, (int)order, ci, da, rtype, (int)batch, s); });
        } else {
            const double* w = wigner_for(plan, (int)order, s);
            const auto fw = rot_work_for(plan, true, (int)order, (int)batch, s);
            timed(plan, ST_ROTATE, (int)batch, s, [&] {
                rotate_fused(zt, (int)order, w, fw.first, fw.second, ci, da, rtype, (int)batch, s);
            });
            const auto work = rot_work_for(plan, false, (int)order, (int)batch, s);
            if (work.second > 0) {
                // degrees above 255: two passes through an intermediate
                ensure(plan->rtemp, plan->rtemp_cap, (size_t)csz * batch);
                for (int pass = 0; pass < 2; ++pass)
                    timed(plan, ST_ROTATE_WIDE, (int)batch, s, [&] {
                        rotate_pass(pass, zt, (int)order, w, work.first, work.second, ci, plan->rtemp, csz,
                                    da, rtype, (int)batch, s);
                    });
            }
        }
        if (mem == ZGL_MEM_HOST) {
            check_hip(hipMemcpy2DAsync(coeffs, cs * sizeof(double), plan->dcoef, csz * sizeof(double),
                                       csz * sizeof(double), batch, hipMemcpyDeviceToHost, s),
                      "coefficient download");
            check_hip(hipStreamSynchronize(s), "rotation");
        }
        return ZGL_SUCCESS;
    });
}
}  // namespace
}  // namespace zgl

extern "C" zgl_status zgl_rotate(zgl_plan* plan, double* coeffs, size_t order, const double* euler_angles,
                                 int rotation_type, size_t batch, ptrdiff_t coeff_stride, zgl_mem mem,
                                 void* stream) {
    return zgl::rotate_entry(plan, coeffs, order, euler_angles, rotation_type, batch, coeff_stride, mem,
                             stream, false);
}

extern "C" zgl_status zgl_polar_rotate(zgl_plan* plan, double* coeffs, size_t order, const double* angles,
                                       int rotation_type, size_t batch, ptrdiff_t coeff_stride, zgl_mem mem,
                                       void* stream) {
    return zgl::rotate_entry(plan, coeffs, order, angles, rotation_type, batch, coeff_stride, mem, stream,
                             true);
}

extern "C" zgl_status zgl_wignerd_pi2(zgl_plan* plan, size_t order, double* out, zgl_mem mem, void* stream) {
    using namespace zgl;
    if (!plan) return fail(ZGL_ERROR_INVALID_VALUE, "plan is NULL");
    if (order > kMaxOrder) return fail(ZGL_ERROR_INVALID_VALUE, "order exceeds the supported maximum");
    if (mem != ZGL_MEM_DEVICE && mem != ZGL_MEM_HOST) return fail(ZGL_ERROR_INVALID_VALUE, "unknown memory kind");
    if (order == 0) return ZGL_SUCCESS;
    if (!out) return fail(ZGL_ERROR_INVALID_VALUE, "NULL pointer");
    if (misaligned(out)) return fail(ZGL_ERROR_INVALID_VALUE, kMisaligned);
    return guarded([&]() -> zgl_status {
        hipStream_t s = static_cast<hipStream_t>(stream);
        const double* w = wigner_for(plan, (int)order, s);
        const size_t n = (size_t)wignerd_size((int)order);
        check_hip(hipMemcpyAsync(out, w, n * sizeof(double),
                                 mem == ZGL_MEM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s),
                  "d table copy");
        if (mem == ZGL_MEM_HOST) check_hip(hipStreamSynchronize(s), "d table copy");
        return ZGL_SUCCESS;
    });
}

extern "C" void zgl_euler_angles_from_rotation_matrix(const double* m, double* angles) {
    if (!m || !angles) return;
    const double r22_sq = m[8] * m[8];
    const double beta_y = std::sqrt((1.0 + r22_sq) * (1.0 - r22_sq));
    angles[0] = std::atan2(m[5], m[2]);
    angles[1] = std::atan2(beta_y, m[8]);
    angles[2] = std::atan2(m[7], -m[6]);
}
