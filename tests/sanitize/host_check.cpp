// Host-side code of the library (the C ABI of include/zgl.h) and the C
// oracle (oracle/zgl_oracle.c) under AddressSanitizer and
// UndefinedBehaviorSanitizer, without a GPU.
//
// Built by tests/sanitize/Makefile: the library's sources host-only (no device
// code: no kernel can run, and none is reached), the oracle and this program
// with -fsanitize=address,undefined; run by tests/test_sanitize.py. Any
// sanitizer report aborts the program with a non-zero status.
//
// Covered:
//  * every entry point of the C ABI with NULL plans and NULL, misaligned or
//    out-of-range arguments: each returns its status before any device work;
//  * order-0 plans of both kinds and every precision: create, resize, sizes,
//    tuning and profiling queries, destroy;
//  * zgl_euler_angles_from_rotation_matrix on random rotations (alpha and
//    gamma of the ZYZ angles that built the matrix come back);
//  * the device-allocator hook: out of memory at the first request, and --
//    with no GPU present -- at every later request of plan creation in turn,
//    each time with every block handed out returned with its size;
//  * the oracle: Zernike and SH round trips over orders, conventions, grid
//    layouts and truncation; evaluation at the quadrature nodes against the
//    backward transform; rotations followed by their inverses, and the power
//    of each degree kept by a rotation; the Gauss-Legendre weights.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "zgl.h"
#include "zgl_oracle.h"

namespace {

int g_failures = 0;

#define CHECK(cond)                                                          \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);    \
            ++g_failures;                                                    \
        }                                                                    \
    } while (0)

std::uint64_t g_rng = 0x9d2c5680a1b2c3d4ull;

double uniform() {  // splitmix64 -> [-1, 1)
    std::uint64_t z = (g_rng += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return double(z >> 11) * 0x1.0p-52 - 1.0;
}

double max_abs(const std::vector<double>& a) {
    double m = 0;
    for (double x : a) m = std::fmax(m, std::fabs(x));
    return m;
}

double max_diff(const std::vector<double>& a, const std::vector<double>& b, size_t n) {
    double m = 0;
    for (size_t i = 0; i < n; ++i) m = std::fmax(m, std::fabs(a[i] - b[i]));
    return m;
}

// ---------------------------------------------------------------- C ABI
void check_invalid_arguments() {
    zgl_plan* p = nullptr;
    CHECK(zgl_plan_create(nullptr, ZGL_ZERNIKE, 0, 0, -1, 0, 4, ZGL_FP64) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_plan_create(&p, (zgl_kind)7, 0, 0, -1, 0, 4, ZGL_FP64) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_plan_create(&p, ZGL_ZERNIKE, 2, 0, -1, 0, 4, ZGL_FP64) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_plan_create(&p, ZGL_ZERNIKE, 0, 3, -1, 0, 4, ZGL_FP64) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_plan_create(&p, ZGL_ZERNIKE, 0, 0, 0, 0, 4, ZGL_FP64) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_plan_create(&p, ZGL_SH, 0, 0, -1, 2, 4, ZGL_FP64) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_plan_create(&p, ZGL_ZERNIKE, 0, 0, -1, 0, 4, (zgl_precision)9) == ZGL_ERROR_NOT_SUPPORTED);
    CHECK(zgl_plan_create(&p, ZGL_ZERNIKE, 0, 0, -1, 0, size_t(1) << 40, ZGL_FP64) == ZGL_ERROR_INVALID_VALUE);
    CHECK(p == nullptr);
    CHECK(std::strlen(zgl_last_error_string()) > 0);

    double buf[64] = {};
    zgl_job job{4, buf, buf};
    CHECK(zgl_plan_resize(nullptr, 4) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_plan_order(nullptr) == 0);
    CHECK(zgl_grid_size(nullptr, 4) == 0 && zgl_coeff_size(nullptr, 4) == 0);
    CHECK(zgl_forward(nullptr, buf, 4, buf, 4, 1, 0, 0, ZGL_MEM_HOST, nullptr) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_backward(nullptr, buf, 4, buf, 4, 1, 0, 0, ZGL_MEM_HOST, nullptr) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_forward_noalloc(nullptr, buf, 4, buf, 4, 1, 0, 0, nullptr) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_forward_ragged(nullptr, &job, 1, nullptr) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_backward_ragged(nullptr, &job, 1, nullptr) == ZGL_ERROR_INVALID_VALUE);
    int n = 0;
    double costs[8];
    CHECK(zgl_ragged_lane_costs(nullptr, costs, 8, &n) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_backward_adjoint(nullptr, buf, 4, buf, 4, 1, 0, 0, ZGL_MEM_HOST, nullptr) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_forward_adjoint(nullptr, buf, 4, buf, 4, 1, 0, 0, ZGL_MEM_HOST, nullptr) == ZGL_ERROR_INVALID_VALUE);
    float fbuf[64] = {};
    CHECK(zgl_forward_grid32(nullptr, fbuf, 4, buf, 4, 1, 0, 0, ZGL_MEM_HOST, nullptr) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_backward_grid32(nullptr, buf, 4, fbuf, 4, 1, 0, 0, ZGL_MEM_HOST, nullptr) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_evaluate(nullptr, buf, 4, buf, 1, buf, 1, buf, 1, buf, 1, 0, 0, ZGL_MEM_HOST, nullptr) ==
          ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_power_spectrum(nullptr, buf, 4, buf, 1, 0, 0, ZGL_MEM_HOST, nullptr) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_cross_power_spectrum(nullptr, buf, buf, 4, buf, 1, 0, 0, ZGL_MEM_HOST, nullptr) ==
          ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_rotate(nullptr, buf, 4, buf, 0, 1, 0, ZGL_MEM_HOST, nullptr) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_polar_rotate(nullptr, buf, 4, buf, 0, 1, 0, ZGL_MEM_HOST, nullptr) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_wignerd_pi2(nullptr, 4, buf, ZGL_MEM_HOST, nullptr) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_plan_autotune(nullptr, 1) == ZGL_ERROR_INVALID_VALUE);
    int entries[10];
    CHECK(zgl_plan_tuning(nullptr, entries, 2, &n) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_profile_enable(nullptr, 1) == ZGL_ERROR_INVALID_VALUE);
    zgl_stage_stats st[4];
    CHECK(zgl_profile_stats(nullptr, st, 4, &n) == ZGL_ERROR_INVALID_VALUE);
    CHECK(zgl_plan_destroy(nullptr) == ZGL_SUCCESS);
}

void check_order0_plans() {
    for (int kind = 0; kind < 2; ++kind)
        for (int prec = 0; prec < 4; ++prec) {
            zgl_plan* p = nullptr;
            const zgl_status s = zgl_plan_create(&p, (zgl_kind)kind, 1, 0, -1, kind, 0, (zgl_precision)prec);
            if (kind == ZGL_SH && prec == ZGL_BF16X3) {
                // BF16x3 is a Zernike mode: either refused or accepted as FP32
                if (s != ZGL_SUCCESS) continue;
            }
            CHECK(s == ZGL_SUCCESS);
            if (!p) continue;
            CHECK(zgl_plan_order(p) == 0);
            CHECK(zgl_plan_resize(p, 0) == ZGL_SUCCESS);
            CHECK(zgl_plan_resize(p, size_t(1) << 30) == ZGL_ERROR_INVALID_VALUE);
            for (size_t N : {0, 1, 2, 7, 8, 33, 128}) {
                const size_t lon = N ? 2 * N - 1 : 0;
                if (kind == ZGL_ZERNIKE) {
                    CHECK(zgl_grid_size(p, N) == (N ? lon * (N + 1) * ((N + 1 + 7) / 8 * 8) : 0));
                    CHECK(zgl_coeff_size(p, N) == (N + 1) * (N + 3) * (2 * N + 1) / 24);
                    CHECK(zgl_grid_size(p, N) == zo_zt_lon_size(N) * zo_zt_lat_size(N) * zo_zt_rad_size(N));
                    CHECK(zgl_coeff_size(p, N) == zo_zt_num_coeffs(N));
                } else {
                    CHECK(zgl_grid_size(p, N) == (N ? zo_st_lon_size(1, N) * zo_st_lat_size(1, N) : 0));
                    CHECK(zgl_coeff_size(p, N) == N * (N + 1) / 2);
                }
            }
            double buf[8] = {};
            // order-0 work is no work; misaligned pointers are refused where used
            CHECK(zgl_forward(p, buf, 0, buf, 0, 3, 0, 0, ZGL_MEM_HOST, nullptr) == ZGL_SUCCESS);
            CHECK(zgl_backward(p, buf, 0, buf, 0, 3, 0, 0, ZGL_MEM_HOST, nullptr) == ZGL_SUCCESS);
            CHECK(zgl_forward(p, buf, 0, buf, 0, 1, 0, 0, (zgl_mem)5, nullptr) == ZGL_ERROR_INVALID_VALUE);
            const double* odd = reinterpret_cast<const double*>(reinterpret_cast<const char*>(buf) + 4);
            const double angles[3] = {0.1, 0.2, 0.3};
            CHECK(zgl_rotate(p, const_cast<double*>(odd), 4, angles, 0, 1, 0, ZGL_MEM_HOST, nullptr) ==
                  ZGL_ERROR_INVALID_VALUE);
            CHECK(std::strstr(zgl_last_error_string(), "aligned") != nullptr);
            CHECK(zgl_wignerd_pi2(p, 3, const_cast<double*>(odd), ZGL_MEM_HOST, nullptr) ==
                  ZGL_ERROR_INVALID_VALUE);
            zgl_job bad{3, const_cast<double*>(odd), buf};
            CHECK(zgl_forward_ragged(p, &bad, 1, nullptr) == ZGL_ERROR_INVALID_VALUE);
            zgl_job null_job{3, nullptr, buf};
            CHECK(zgl_forward_ragged(p, &null_job, 1, nullptr) == ZGL_ERROR_INVALID_VALUE);
            CHECK(zgl_forward_ragged(p, nullptr, 2, nullptr) == ZGL_ERROR_INVALID_VALUE);
            CHECK(zgl_forward_ragged(p, nullptr, 0, nullptr) == ZGL_SUCCESS);
            CHECK(zgl_plan_autotune(p, 1) == ZGL_SUCCESS);
            int n = -1;
            int entries[20];
            CHECK(zgl_plan_tuning(p, entries, 4, &n) == ZGL_SUCCESS && n == 0);
            CHECK(zgl_profile_enable(p, 1) == ZGL_SUCCESS);
            zgl_stage_stats st[16];
            CHECK(zgl_profile_stats(p, st, 16, &n) == ZGL_SUCCESS);
            CHECK(zgl_profile_enable(p, 0) == ZGL_SUCCESS);
            CHECK(zgl_plan_destroy(p) == ZGL_SUCCESS);
        }
}

void matmul3(const double* a, const double* b, double* c) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            c[3 * i + j] = 0;
            for (int k = 0; k < 3; ++k) c[3 * i + j] += a[3 * i + k] * b[3 * k + j];
        }
}

void rot_z(double t, double* m) {
    const double c = std::cos(t), s = std::sin(t);
    const double r[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
    std::memcpy(m, r, sizeof r);
}

void rot_y(double t, double* m) {
    const double c = std::cos(t), s = std::sin(t);
    const double r[9] = {c, 0, s, 0, 1, 0, -s, 0, c};
    std::memcpy(m, r, sizeof r);
}

void check_euler_angles() {
    for (int trial = 0; trial < 200; ++trial) {
        const double a = 3.0 * uniform(), b = 1.5 * (uniform() + 1.0) + 0.01, g = 3.0 * uniform();
        double za[9], yb[9], zg[9], t[9], r[9];
        rot_z(a, za);
        rot_y(b, yb);
        rot_z(g, zg);
        matmul3(za, yb, t);
        matmul3(t, zg, r);
        double ang[3];
        zgl_euler_angles_from_rotation_matrix(r, ang);
        // alpha and gamma rebuild the matrix's third row and column; beta is
        // zest's expression, atan2(sqrt(1 - r22^4), r22)
        // (include/rotor.hpp:55-59), compared with zest itself in
        // tests/test_abi.py
        CHECK(std::fabs(std::remainder(ang[0] - a, 2.0 * M_PI)) < 1e-12);
        CHECK(std::fabs(std::remainder(ang[2] - g, 2.0 * M_PI)) < 1e-12);
        CHECK(ang[1] >= 0.0 && ang[1] <= M_PI);
    }
}

// Host memory handed out as if it were device memory: only valid while no
// GPU is present (no kernel can then run on it).
struct HostPool {
    size_t allow = 0;  // requests that succeed
    size_t calls = 0;
    std::map<void*, std::pair<void*, size_t>> live;  // aligned -> (raw, bytes)
};

void* pool_alloc(size_t bytes, void* user) {
    auto* h = static_cast<HostPool*>(user);
    if (h->calls++ >= h->allow) return nullptr;
    void* raw = std::malloc(bytes + 256);
    if (!raw) return nullptr;
    void* p = reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(raw) + 255) & ~uintptr_t(255));
    h->live[p] = {raw, bytes};
    return p;
}

void pool_free(void* p, size_t bytes, void* user) {
    auto* h = static_cast<HostPool*>(user);
    auto it = h->live.find(p);
    if (it == h->live.end() || it->second.second != bytes) {
        std::printf("FAILED: free of a block not handed out (or with another size)\n");
        ++g_failures;
        return;
    }
    std::free(it->second.first);
    h->live.erase(it);
}

void check_allocator_hook() {
    CHECK(zgl_set_device_allocator(pool_alloc, nullptr, nullptr) == ZGL_ERROR_INVALID_VALUE);
    HostPool none;
    CHECK(zgl_set_device_allocator(pool_alloc, pool_free, &none) == ZGL_SUCCESS);
    zgl_plan* p = nullptr;
    CHECK(zgl_plan_create(&p, ZGL_ZERNIKE, 1, 0, -1, 0, 8, ZGL_FP64) == ZGL_ERROR_OUT_OF_MEMORY);
    CHECK(p == nullptr && none.calls == 1 && none.live.empty());
    CHECK(std::strstr(zgl_last_error_string(), "returned NULL") != nullptr);
    CHECK(zgl_set_device_allocator(nullptr, nullptr, nullptr) == ZGL_SUCCESS);

    int devices = 0;
    if (hipGetDeviceCount(&devices) == hipSuccess && devices > 0) {
        std::printf("a GPU is present: later-request injection skipped\n");
        return;
    }
    struct Case {
        zgl_kind kind;
        size_t order;
        zgl_precision prec;
    };
    for (const Case c : {Case{ZGL_ZERNIKE, 12, ZGL_FP64}, Case{ZGL_ZERNIKE, 12, ZGL_FP32},
                         Case{ZGL_SH, 16, ZGL_FP64}, Case{ZGL_SH, 9, ZGL_FP32}}) {
        size_t k = 0;
        for (;; ++k) {
            HostPool pool;
            pool.allow = k;
            CHECK(zgl_set_device_allocator(pool_alloc, pool_free, &pool) == ZGL_SUCCESS);
            zgl_plan* q = nullptr;
            const zgl_status s = zgl_plan_create(&q, c.kind, 1, 0, -1, 0, c.order, c.prec);
            CHECK(zgl_set_device_allocator(nullptr, nullptr, nullptr) == ZGL_SUCCESS);
            CHECK(q == nullptr);
            CHECK(pool.live.empty());
            if (s != ZGL_ERROR_OUT_OF_MEMORY) {
                CHECK(s == ZGL_ERROR_HIP);  // the first checked kernel, with no device
                break;
            }
            CHECK(pool.calls == k + 1);
            if (k > 100) {
                CHECK(false);
                break;
            }
        }
        CHECK(k >= 3);
        std::printf("allocator: kind %d order %zu precision %d: %zu allocations before the first kernel\n",
                    (int)c.kind, c.order, (int)c.prec, k);
    }
}

// ---------------------------------------------------------------- oracle
std::vector<double> zernike_coeffs(size_t N) {
    std::vector<double> c(2 * zo_zt_num_coeffs(N));
    for (size_t n = 0; n < N; ++n)
        for (size_t l = n & 1; l <= n; l += 2)
            for (size_t m = 0; m <= l; ++m) {
                const size_t i = (n + 1) * (n + 3) * (2 * n + 1) / 24 + l * l / 4 + m;
                c[2 * i] = uniform();
                c[2 * i + 1] = m ? uniform() : 0.0;
            }
    return c;
}

std::vector<double> sh_coeffs(size_t L) {
    std::vector<double> c(L * (L + 1));
    for (size_t l = 0; l < L; ++l)
        for (size_t m = 0; m <= l; ++m) {
            const size_t i = l * (l + 1) / 2 + m;
            c[2 * i] = uniform();
            c[2 * i + 1] = m ? uniform() : 0.0;
        }
    return c;
}

void check_oracle_zernike() {
    for (size_t N : {1, 2, 3, 5, 8, 9, 16, 33}) {
        for (int conv = 0; conv < 8; ++conv) {
            const int zn = conv >> 2, sn = (conv >> 1) & 1, sp = (conv & 1) ? 1 : -1;
            const auto c = zernike_coeffs(N);
            std::vector<double> g(zo_zt_lon_size(N) * zo_zt_lat_size(N) * zo_zt_rad_size(N));
            CHECK(zo_zt_backward(zn, sn, sp, N, c.data(), N, g.data()) == 0);
            std::vector<double> c2(c.size(), 0.0);
            CHECK(zo_zt_forward(zn, sn, sp, N, g.data(), N, c2.data()) == 0);
            CHECK(max_diff(c, c2, c.size()) < 1e-12 * max_abs(c));
        }
        // truncation both ways
        const size_t M = N + 3;
        const auto c = zernike_coeffs(N);
        std::vector<double> gM(zo_zt_lon_size(M) * zo_zt_lat_size(M) * zo_zt_rad_size(M));
        CHECK(zo_zt_backward(1, 0, -1, N, c.data(), M, gM.data()) == 0);
        std::vector<double> cM(2 * zo_zt_num_coeffs(M), 7.0);
        CHECK(zo_zt_forward(1, 0, -1, M, gM.data(), M, cM.data()) == 0);
        CHECK(max_diff(c, cM, c.size()) < 1e-12 * max_abs(c));
        std::vector<double> tail(cM.begin() + c.size(), cM.end());
        CHECK(max_abs(tail) < 1e-12 * max_abs(c));
        // evaluation at the quadrature nodes equals the backward grid
        if (N <= 9) {
            // the radial nodes fill the padded radial extent of the grid
            const size_t nphi = zo_zt_lon_size(N), nth = zo_zt_lat_size(N), nr = zo_zt_rad_size(N);
            std::vector<double> lon(nphi), z(nth), wz(nth), x(nr), wx(nr), colat(nth), r(nr);
            for (size_t i = 0; i < nphi; ++i) lon[i] = 2.0 * M_PI * double(i) / double(nphi);
            zo_gl_nodes_weights(nth, z.data(), wz.data());
            zo_gl_nodes_weights(nr, x.data(), wx.data());
            for (size_t i = 0; i < nth; ++i) colat[i] = std::acos(z[i]);
            for (size_t i = 0; i < nr; ++i) r[i] = 0.5 * (1.0 + x[i]);
            std::vector<double> v(nphi * nth * nr);
            zo_zt_evaluate(1, 0, -1, N, c.data(), lon.data(), nphi, colat.data(), nth, r.data(), nr, v.data());
            std::vector<double> g(zo_zt_lon_size(N) * zo_zt_lat_size(N) * zo_zt_rad_size(N));
            CHECK(zo_zt_backward(1, 0, -1, N, c.data(), N, g.data()) == 0);
            CHECK(max_diff(v, g, g.size()) < 1e-12 * max_abs(g));
        }
    }
}

void check_oracle_sh() {
    for (size_t L : {1, 2, 3, 7, 8, 17, 40}) {
        for (int conv = 0; conv < 4; ++conv) {
            const int sn = conv >> 1, sp = (conv & 1) ? 1 : -1;
            const auto c = sh_coeffs(L);
            std::vector<double> g(zo_st_lon_size(0, L) * zo_st_lat_size(0, L));
            CHECK(zo_st_backward(sn, sp, 0, L, c.data(), L, g.data()) == 0);
            std::vector<double> c2(c.size());
            CHECK(zo_st_forward(sn, sp, 0, L, g.data(), L, c2.data()) == 0);
            CHECK(max_diff(c, c2, c.size()) < 1e-12 * max_abs(c));
            // LatLon (its own node counts: the padded longitude extent is
            // sampled). zest's LatLon forward does not invert its backward at
            // odd orders; the oracle follows zest there, so the round trip is
            // checked at even orders only.
            std::vector<double> h(zo_st_lon_size(1, L) * zo_st_lat_size(1, L));
            CHECK(zo_st_backward(sn, sp, 1, L, c.data(), L, h.data()) == 0);
            CHECK(zo_st_forward(sn, sp, 1, L, h.data(), L, c2.data()) == 0);
            if (L % 2 == 0) CHECK(max_diff(c, c2, c.size()) < 1e-12 * max_abs(c));
        }
        // a rotation keeps the power of every degree; its inverse restores
        // the coefficients
        const auto c = sh_coeffs(L);
        auto r = c;
        zo_rotate(1, L, r.data(), 0.4, 1.1, -0.7, 0);
        std::vector<double> ps(L), pr(L);
        zo_st_power_spectrum(1, L, c.data(), c.data(), ps.data());
        zo_st_power_spectrum(1, L, r.data(), r.data(), pr.data());
        CHECK(max_diff(ps, pr, L) < 1e-12 * max_abs(ps));
        zo_rotate(1, L, r.data(), 0.7, -1.1, -0.4, 0);
        CHECK(max_diff(r, c, c.size()) < 1e-12 * max_abs(c));
        auto q = c;
        zo_polar_rotate(1, L, q.data(), 0.9, 0);
        zo_polar_rotate(1, L, q.data(), -0.9, 0);
        CHECK(max_diff(q, c, c.size()) < 1e-12 * max_abs(c));
    }
    for (size_t n : {1, 2, 3, 10, 31, 64, 257}) {
        std::vector<double> z(n), w(n);
        zo_gl_nodes_weights(n, z.data(), w.data());
        double s = 0;
        for (double x : w) s += x;
        CHECK(std::fabs(s - 2.0) < 1e-13);
    }
}

}  // namespace

// A read one past the end of a heap block, run only when asked for
// (--planted-overflow): shows that the instrumentation of this build reports.
int planted_overflow() {
    volatile int n = 8;
    int* v = static_cast<int*>(std::malloc(sizeof(int) * n));
    for (int i = 0; i < n; ++i) v[i] = i;
    const int x = v[n];
    std::free(v);
    return x;
}

int main(int argc, char** argv) {
    if (argc > 1 && std::strcmp(argv[1], "--planted-overflow") == 0) return planted_overflow() == 12345;
    check_invalid_arguments();
    check_order0_plans();
    check_euler_angles();
    check_allocator_hook();
    check_oracle_zernike();
    check_oracle_sh();
    std::printf("%d failures\n", g_failures);
    return g_failures ? 1 : 0;
}
