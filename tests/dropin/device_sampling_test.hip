// Device-side sampling (include/zest_amd/device_sampling.hpp) against zest's
// host-side samplers (BallGLQGridPoints / SphereGLQGridPoints::generate_values
// of the drop-in headers, which keep zest's coordinates): the grids agree to
// a few ulp (device and host sin/cos/exp), and so do their forward transforms.
// Exit status 0 on success; prints the largest differences.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstdio>
#include <vector>

#include "zernike_glq_transformer.hpp"
#include "sh_glq_transformer.hpp"
#include "zest_amd/device_sampling.hpp"

struct BallFunction {
    __host__ __device__ double operator()(int b, double lon, double colat, double r) const {
        const double s = sin(colat);
        const double x = r * s * cos(lon), y = r * s * sin(lon), z = r * cos(colat);
        return (1.0 + 0.25 * b) * r * exp(-(x - 0.3) * (x - 0.3) - 2.0 * y * y) + z * z * z;
    }
};

struct SphereFunction {
    __host__ __device__ double operator()(int b, double lon, double colat) const {
        return cos(3.0 * lon + 0.1 * b) * sin(colat) * sin(colat) + exp(cos(colat));
    }
};

static double max_rel(const std::vector<double>& a, const std::vector<double>& b) {
    double d = 0.0, m = 0.0;
    for (std::size_t i = 0; i < a.size(); ++i) {
        d = std::max(d, std::fabs(a[i] - b[i]));
        m = std::max(m, std::fabs(b[i]));
    }
    return m > 0 ? d / m : d;
}

#define CHECK(x)                                                             \
    do {                                                                     \
        if ((x) != 0) {                                                      \
            std::printf("failed: %s (%s)\n", #x, zgl_last_error_string());  \
            return 1;                                                        \
        }                                                                    \
    } while (0)

int main() {
    bool ok = true;
    // ---- ball, order 32, 3 functions
    {
        constexpr std::size_t order = 32, batch = 3;
        using Layout = zest::zt::LonLatRadLayout<>;
        zest::zt::BallGLQGridPoints<Layout> points(order);
        const auto shape = Layout::shape(order);
        const std::size_t n = shape[0] * shape[1] * shape[2];
        std::vector<double> host(n * batch), dev(n * batch);
        for (std::size_t b = 0; b < batch; ++b) {
            zest::zt::BallGLQGrid<double, Layout> g(order);
            points.generate_values(g, [&](double lon, double colat, double r) {
                return BallFunction{}(int(b), lon, colat, r);
            });
            std::copy(g.flatten().begin(), g.flatten().end(), host.begin() + b * n);
        }
        double* d = nullptr;
        CHECK(hipMalloc(reinterpret_cast<void**>(&d), n * batch * sizeof(double)));
        zest_amd::sample_ball_glq_grid(order, batch, BallFunction{}, d, 0, nullptr);
        CHECK(hipMemcpy(dev.data(), d, n * batch * sizeof(double), hipMemcpyDeviceToHost));
        const double eg = max_rel(dev, host);
        // forward transforms of both grids
        zgl_plan* p = nullptr;
        CHECK(zgl_plan_create(&p, ZGL_ZERNIKE, ZGL_ZNORM_NORMED, ZGL_SHNORM_QM, ZGL_PHASE_NONE,
                              ZGL_LAYOUT_LONLAT, order, ZGL_FP64));
        const std::size_t nc = 2 * zgl_coeff_size(p, order);
        std::vector<double> ch(nc * batch), cd(nc * batch);
        CHECK(zgl_forward(p, host.data(), order, ch.data(), order, batch, 0, 0, ZGL_MEM_HOST, nullptr));
        double* dc = nullptr;
        CHECK(hipMalloc(reinterpret_cast<void**>(&dc), nc * batch * sizeof(double)));
        CHECK(zgl_forward(p, d, order, dc, order, batch, 0, 0, ZGL_MEM_DEVICE, nullptr));
        CHECK(hipMemcpy(cd.data(), dc, nc * batch * sizeof(double), hipMemcpyDeviceToHost));
        const double ec = max_rel(cd, ch);
        std::printf("ball order %zu: grid %.3e, coefficients %.3e\n", order, eg, ec);
        ok = ok && eg < 1e-14 && ec < 1e-13;
        zgl_plan_destroy(p);
        CHECK(hipFree(d));
        CHECK(hipFree(dc));
    }
    // ---- sphere, order 40, both layouts, 2 functions
    for (int layout : {ZGL_LAYOUT_LONLAT, ZGL_LAYOUT_LATLON}) {
        constexpr std::size_t order = 40, batch = 2;
        std::vector<double> host, dev;
        std::size_t n = 0;
        auto fill = [&](auto points, auto grid_of) {
            for (std::size_t b = 0; b < batch; ++b) {
                auto g = grid_of();
                points.generate_values(g, [&](double lon, double colat) { return SphereFunction{}(int(b), lon, colat); });
                n = g.flatten().size();
                host.insert(host.end(), g.flatten().begin(), g.flatten().end());
            }
        };
        if (layout == ZGL_LAYOUT_LONLAT)
            fill(zest::st::SphereGLQGridPoints<zest::st::LonLatLayout<>>(order),
                 [] { return zest::st::SphereGLQGrid<double, zest::st::LonLatLayout<>>(order); });
        else
            fill(zest::st::SphereGLQGridPoints<zest::st::LatLonLayout<>>(order),
                 [] { return zest::st::SphereGLQGrid<double, zest::st::LatLonLayout<>>(order); });
        dev.resize(host.size());
        double* d = nullptr;
        CHECK(hipMalloc(reinterpret_cast<void**>(&d), host.size() * sizeof(double)));
        zest_amd::sample_sphere_glq_grid(order, layout, batch, SphereFunction{}, d, 0, nullptr);
        CHECK(hipMemcpy(dev.data(), d, host.size() * sizeof(double), hipMemcpyDeviceToHost));
        const double eg = max_rel(dev, host);
        std::printf("sphere order %zu layout %d: grid %.3e (%zu points per instance)\n", order, layout, eg, n);
        ok = ok && eg < 1e-14;
        CHECK(hipFree(d));
    }
    std::printf(ok ? "ok\n" : "FAILED\n");
    return ok ? 0 : 1;
}
