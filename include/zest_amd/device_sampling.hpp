// Sampling of functions on the quadrature grids on the GPU, for callers
// compiled with hipcc.
//
// zest's grid-point samplers (BallGLQGridPoints::generate_values,
// include/zernike_glq_transformer.hpp:397-409; SphereGLQGridPoints::
// generate_values, include/sh_glq_transformer.hpp:433-448) and the
// transformers built on them (ZernikeTransformer, SHTransformer) evaluate the
// function at every grid point on the CPU; the grid then has to reach the GPU
// (36 MB per instance at order 128). Here the function is a __device__
// callable evaluated by a kernel at the same coordinates -- computed on the
// host with zest's own Gauss-Legendre routines, exactly as the samplers
// compute them -- and written in zest's grid layout straight into device
// memory, where zgl_forward reads it. Only the coordinate arrays cross PCIe.
//
//   sample_ball_glq_grid(order, batch, f, d_grid, grid_stride, stream)
//       d_grid[b][i][j][k] = f(b, lon_i, colat_j, r_k)   (LonLatRadLayout)
//   sample_sphere_glq_grid(order, layout, batch, f, d_grid, grid_stride, stream)
//       f(b, lon, colat) on LonLatLayout ([lon][lat]) or LatLonLayout ([lat][lon])
//
// grid_stride is the distance in doubles between instances (0: packed). The
// call enqueues the kernel on `stream` and returns.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <numbers>
#include <stdexcept>
#include <vector>

#include "gauss_legendre.hpp"
#include "zgl.h"

namespace zest_amd {

namespace detail_sampling {

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// lon, colat (and r) of the quadrature grid, uploaded to one device buffer
struct Coordinates {
    double* d = nullptr;
    int nlon = 0, nlat = 0, nrad = 0;
    Coordinates(int nlon_, int nlat_, int nrad_) : nlon(nlon_), nlat(nlat_), nrad(nrad_) {
        std::vector<double> h(std::size_t(nlon + nlat + nrad));
        const double dlon = (2.0 * std::numbers::pi) / double(nlon);
        for (int i = 0; i < nlon; ++i) h[i] = dlon * double(i);
        std::span<double> lat(h.data() + nlon, std::size_t(nlat));
        zest::gl::gl_nodes<zest::gl::UnpackedLayout, zest::gl::GLNodeStyle::angle>(lat, nlat & 1);
        if (nrad > 0) {
            std::span<double> rad(h.data() + nlon + nlat, std::size_t(nrad));
            zest::gl::gl_nodes<zest::gl::UnpackedLayout, zest::gl::GLNodeStyle::cos>(rad, nrad & 1);
            for (auto& x : rad) x = 0.5 * (1.0 + x);
        }
        check(hipMalloc(reinterpret_cast<void**>(&d), h.size() * sizeof(double)), "coordinate buffer");
        check(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice), "coordinate upload");
    }
    ~Coordinates() {
        if (d) (void)hipFree(d);  // waits for the sampling kernel
    }
    Coordinates(const Coordinates&) = delete;
    Coordinates& operator=(const Coordinates&) = delete;
};

template <class F>
__global__ void ball_kernel(F f, const double* lon, const double* colat, const double* r, int nlon, int nlat,
                            int nrad, double* grid, long stride) {
    const long n = (long)nlon * nlat * nrad;
    const int b = blockIdx.y;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const int k = (int)(e % nrad);
        const long ij = e / nrad;
        const int j = (int)(ij % nlat), i = (int)(ij / nlat);
        grid[b * stride + e] = f(b, lon[i], colat[j], r[k]);
    }
}

template <class F>
__global__ void sphere_kernel(F f, const double* lon, const double* colat, int nlon, int nlat, int latlon,
                              double* grid, long stride) {
    const long n = (long)nlon * nlat;
    const int b = blockIdx.y;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        int i, j;  // longitude, latitude
        if (latlon) {
            j = (int)(e / nlon);
            i = (int)(e % nlon);
        } else {
            i = (int)(e / nlat);
            j = (int)(e % nlat);
        }
        grid[b * stride + e] = f(b, lon[i], colat[j]);
    }
}

inline unsigned blocks_for(long n) {
    const long b = (n + 255) / 256;
    return (unsigned)(b < 4096 ? b : 4096);
}

}  // namespace detail_sampling

// Ball grid of `order` (zest's BallGLQGrid, LonLatRadLayout: 2N-1 longitudes,
// N+1 colatitudes, round8(N+1) radii) for `batch` functions f(b, lon, colat, r).
template <class F>
void sample_ball_glq_grid(std::size_t order, std::size_t batch, F f, double* d_grid, std::ptrdiff_t grid_stride,
                          hipStream_t stream) {
    if (order == 0 || batch == 0) return;
    if (batch > 65535) throw std::invalid_argument("batch exceeds 65535");
    const int nlon = int(2 * order - 1), nlat = int(order + 1), nrad = int((order + 1 + 7) / 8 * 8);
    const long n = (long)nlon * nlat * nrad;
    const long stride = grid_stride ? (long)grid_stride : n;
    detail_sampling::Coordinates c(nlon, nlat, nrad);
    detail_sampling::ball_kernel<<<dim3(detail_sampling::blocks_for(n), unsigned(batch)), 256, 0, stream>>>(
        f, c.d, c.d + nlon, c.d + nlon + nlat, nlon, nlat, nrad, d_grid, stride);
    detail_sampling::check(hipGetLastError(), "ball sampling launch");
}

// Sphere grid of `order` (zest's SphereGLQGrid) in ZGL_LAYOUT_LONLAT (2L-1
// longitudes x round8(L) colatitudes) or ZGL_LAYOUT_LATLON (L colatitudes x
// round8(2L-1) longitudes) for `batch` functions f(b, lon, colat).
template <class F>
void sample_sphere_glq_grid(std::size_t order, int layout, std::size_t batch, F f, double* d_grid,
                            std::ptrdiff_t grid_stride, hipStream_t stream) {
    if (order == 0 || batch == 0) return;
    if (batch > 65535) throw std::invalid_argument("batch exceeds 65535");
    const bool latlon = layout == ZGL_LAYOUT_LATLON;
    const int nlon = latlon ? int((2 * order - 1 + 7) / 8 * 8) : int(2 * order - 1);
    const int nlat = latlon ? int(order) : int((order + 7) / 8 * 8);
    const long n = (long)nlon * nlat;
    const long stride = grid_stride ? (long)grid_stride : n;
    detail_sampling::Coordinates c(nlon, nlat, 0);
    detail_sampling::sphere_kernel<<<dim3(detail_sampling::blocks_for(n), unsigned(batch)), 256, 0, stream>>>(
        f, c.d, c.d + nlon, nlon, nlat, latlon ? 1 : 0, d_grid, stride);
    detail_sampling::check(hipGetLastError(), "sphere sampling launch");
}

}  // namespace zest_amd
