// Drop-in replacement of zest's "zernike_glq_transformer.hpp" (sebsassi/zest,
// include/zernike_glq_transformer.hpp) backed by the MI355X library
// (include/zgl.h). Put this directory ahead of zest's include directory, or
// include "zest_amd/zernike_glq_transformer.hpp" instead of zest's header: the
// names, namespaces, template parameters, memory layouts and convention types
// are zest's, so caller code compiles unchanged. zest's own data-model
// headers (expansions, layouts, conventions, MDSpan, Gauss-Legendre helpers)
// are used as they are.
//
// What runs where: GLQTransformer::forward_transform / backward_transform
// (zest: 640-707) call zgl_forward / zgl_backward, which stage the caller's
// host buffers through the GPU and return when the output is written, like
// zest. The grid views, the grid container and the grid-point sampler are
// host-side data types and keep zest's behaviour.
//
// Supported grid layout: LonLatRadLayout<CacheLineAlignment> (zest's default,
// radial size padded to a multiple of 8).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <concepts>
#include <cstddef>
#include <numbers>
#include <span>
#include <type_traits>
#include <vector>

#include "alignment.hpp"
#include "gauss_legendre.hpp"
#include "md_span.hpp"
#include "plm_recursion.hpp"
#include "radial_zernike_recursion.hpp"
#include "zernike_expansion.hpp"

#include "zgl_plan.hpp"

namespace zest {
namespace zt {

// Grid layout [lon][lat][rad] (zest: include/zernike_glq_transformer.hpp:50-131).
template <typename AlignmentType = CacheLineAlignment>
struct LonLatRadLayout {
    using Alignment = AlignmentType;

    [[nodiscard]] static constexpr std::size_t lon_size(std::size_t order) noexcept {
        return order ? 2 * order - 1 : 0;
    }
    [[nodiscard]] static constexpr std::size_t lat_size(std::size_t order) noexcept { return order + 1; }
    [[nodiscard]] static constexpr std::size_t rad_size(std::size_t order) noexcept {
        if constexpr (std::is_same_v<Alignment, NoAlignment>)
            return order + 1;
        else
            return zest::detail::next_divisible<Alignment::template vector_size<double>()>(order + 1);
    }
    [[nodiscard]] static constexpr std::size_t size(std::size_t order) noexcept {
        return lon_size(order) * lat_size(order) * rad_size(order);
    }
    [[nodiscard]] static constexpr std::array<std::size_t, 3> shape(std::size_t order) noexcept {
        return {lon_size(order), lat_size(order), rad_size(order)};
    }
    [[nodiscard]] static constexpr std::size_t fft_size(std::size_t order) noexcept {
        return (lon_size(order) >> 1) + 1;
    }
    [[nodiscard]] static constexpr std::array<std::size_t, 3> fft_stride(std::size_t order) noexcept {
        return {lat_size(order) * rad_size(order), rad_size(order), 1};
    }

    static constexpr std::size_t lat_axis = 1;
    static constexpr std::size_t lon_axis = 0;
    static constexpr std::size_t rad_axis = 2;
};

using DefaultLayout = LonLatRadLayout<>;

// Non-owning view of a ball grid (zest: 141-213).
template <typename ElementType, typename LayoutType = DefaultLayout>
class BallGLQGridSpan : public MDSpan<ElementType, 3> {
public:
    using Base = MDSpan<ElementType, 3>;
    using typename Base::element_type;
    using Layout = LayoutType;
    using ConstView = BallGLQGridSpan<const element_type, Layout>;
    using Base::data;
    using Base::extents;
    using Base::size;

    [[nodiscard]] static constexpr std::size_t size(std::size_t order) noexcept { return Layout::size(order); }
    [[nodiscard]] static constexpr std::array<std::size_t, 3> shape(std::size_t order) noexcept {
        return Layout::shape(order);
    }

    BallGLQGridSpan() noexcept = default;
    constexpr BallGLQGridSpan(element_type* data, std::size_t order) noexcept
        : Base(data, Layout::shape(order)), m_order(order) {}
    constexpr BallGLQGridSpan(std::span<element_type> buffer, std::size_t order) noexcept
        : Base(buffer.data(), Layout::shape(order)), m_order(order) {}

    [[nodiscard]] constexpr std::size_t order() const noexcept { return m_order; }
    [[nodiscard]] constexpr const std::array<std::size_t, 3>& shape() const noexcept { return extents(); }
    [[nodiscard]] constexpr std::span<element_type> flatten() const noexcept {
        return std::span<element_type>(data(), size());
    }
    // Through MDSpan's (data, size, extents) constructor, which stays usable
    // in constant expressions.
    [[nodiscard]] constexpr operator ConstView() const noexcept {
        return ConstView(data(), size(), extents(), m_order);
    }

private:
    friend BallGLQGridSpan<std::remove_const_t<element_type>, Layout>;
    constexpr BallGLQGridSpan(element_type* data, std::size_t size, const std::array<std::size_t, 3>& extents,
                    std::size_t order) noexcept
        : Base(data, size, extents), m_order(order) {}

    std::size_t m_order{};
};

// Owning ball grid (zest: 221-315).
template <typename ElementType, typename LayoutType = DefaultLayout>
class BallGLQGrid {
public:
    using element_type = ElementType;
    using value_type = std::remove_cvref_t<element_type>;
    using Layout = LayoutType;
    using View = BallGLQGridSpan<double>;
    using ConstView = BallGLQGridSpan<const double>;

    [[nodiscard]] static constexpr std::size_t size(std::size_t order) noexcept { return Layout::size(order); }
    [[nodiscard]] static constexpr std::array<std::size_t, 3> shape(std::size_t order) noexcept {
        return Layout::shape(order);
    }

    BallGLQGrid() = default;
    explicit BallGLQGrid(std::size_t order)
        : m_data(Layout::size(order)), m_shape(Layout::shape(order)), m_order(order) {}

    [[nodiscard]] std::size_t order() const noexcept { return m_order; }
    [[nodiscard]] std::array<std::size_t, 3> shape() { return m_shape; }
    [[nodiscard]] std::span<const element_type> flatten() const noexcept { return m_data; }
    std::span<element_type> flatten() noexcept { return m_data; }

    [[nodiscard]] operator View() noexcept { return View(m_data, m_order); }
    [[nodiscard]] operator ConstView() const noexcept { return ConstView(m_data, m_order); }

    void resize(std::size_t order) {
        m_data.resize(Layout::size(order));
        m_shape = Layout::shape(order);
        m_order = order;
    }

    [[nodiscard]] element_type operator()(std::size_t i, std::size_t j, std::size_t k) const noexcept {
        return m_data[(i * m_shape[1] + j) * m_shape[2] + k];
    }
    [[nodiscard]] element_type& operator()(std::size_t i, std::size_t j, std::size_t k) noexcept {
        return m_data[(i * m_shape[1] + j) * m_shape[2] + k];
    }

private:
    std::vector<element_type, AlignedAllocator<element_type, typename LayoutType::Alignment>> m_data{};
    std::array<std::size_t, 3> m_shape{};
    std::size_t m_order{};
};

template <typename T>
concept ball_glq_grid =
    std::same_as<std::remove_cvref_t<T>, BallGLQGridSpan<typename std::remove_cvref_t<T>::element_type,
                                                          typename std::remove_cvref_t<T>::Layout>> ||
    std::same_as<std::remove_cvref_t<T>, BallGLQGrid<typename std::remove_cvref_t<T>::element_type,
                                                      typename std::remove_cvref_t<T>::Layout>>;

// Grid sample points and function sampling (zest: 335-501): longitudes
// 2 pi i / nlon, colatitudes at the Gauss-Legendre angles, radii (1 + x) / 2.
template <typename LayoutType = DefaultLayout>
class BallGLQGridPoints {
public:
    using GridLayout = LayoutType;

    BallGLQGridPoints() = default;
    explicit BallGLQGridPoints(std::size_t order) { resize(order); }

    void resize(std::size_t order) {
        const auto s = GridLayout::shape(order);
        resize(s[GridLayout::lon_axis], s[GridLayout::lat_axis], s[GridLayout::rad_axis]);
    }

    [[nodiscard]] std::span<const double> longitudes() const noexcept { return m_lon; }
    [[nodiscard]] std::span<const double> rad_glq_nodes() const noexcept { return m_rad; }
    [[nodiscard]] std::span<const double> lat_glq_nodes() const noexcept { return m_lat; }

    template <ball_glq_grid GridType, typename FuncType>
        requires std::same_as<typename std::remove_cvref_t<GridType>::Layout, GridLayout>
    void generate_values(GridType&& grid, FuncType&& f) {
        resize(grid.order());
        for (std::size_t i = 0; i < m_lon.size(); ++i)
            for (std::size_t j = 0; j < m_lat.size(); ++j)
                for (std::size_t k = 0; k < m_rad.size(); ++k) grid(i, j, k) = f(m_lon[i], m_lat[j], m_rad[k]);
    }

    template <typename FuncType>
    [[nodiscard]] auto generate_values(FuncType&& f, std::size_t order) {
        using CodomainType = std::invoke_result_t<FuncType, double, double, double>;
        BallGLQGrid<CodomainType, GridLayout> grid(order);
        generate_values(grid, f);
        return grid;
    }

private:
    void resize(std::size_t nlon, std::size_t nlat, std::size_t nrad) {
        if (nlon != m_lon.size()) {
            m_lon.resize(nlon);
            const double dlon = (2.0 * std::numbers::pi) / double(nlon);
            for (std::size_t i = 0; i < nlon; ++i) m_lon[i] = dlon * double(i);
        }
        if (nlat != m_lat.size()) {
            m_lat.resize(nlat);
            gl::gl_nodes<gl::UnpackedLayout, gl::GLNodeStyle::angle>(m_lat, nlat & 1);
        }
        if (nrad != m_rad.size()) {
            m_rad.resize(nrad);
            gl::gl_nodes<gl::UnpackedLayout, gl::GLNodeStyle::cos>(m_rad, nrad & 1);
            for (auto& x : m_rad) x = 0.5 * (1.0 + x);
        }
    }

    std::vector<double> m_rad{};
    std::vector<double> m_lat{};
    std::vector<double> m_lon{};
};

// Ball grid <-> real Zernike expansion (zest: 511-965), on the GPU.
template <ZernikeNorm zernike_norm_param, st::SHNorm sh_norm_param, st::SHPhase sh_phase_param,
          typename GridLayoutType = DefaultLayout>
class GLQTransformer {
    static_assert(std::same_as<GridLayoutType, LonLatRadLayout<CacheLineAlignment>>,
                  "the MI355X transformer supports the default LonLatRadLayout");

public:
    using GridLayout = GridLayoutType;
    static constexpr ZernikeNorm zernike_norm = zernike_norm_param;
    static constexpr st::SHNorm sh_norm = sh_norm_param;
    static constexpr st::SHPhase phase = sh_phase_param;

    GLQTransformer() : m_plan(make_plan(0)) {}
    explicit GLQTransformer(std::size_t order) : m_plan(make_plan(order)) {}

    [[nodiscard]] std::size_t order() const noexcept { return m_plan.order(); }
    void resize(std::size_t order) { m_plan.resize(order); }

    void forward_transform(
        BallGLQGridSpan<const double, GridLayout> values,
        RealZernikeSpan<std::array<double, 2>, zernike_norm_param, sh_norm_param, sh_phase_param> expansion) {
        m_plan.forward(values.data(), values.order(), reinterpret_cast<double*>(expansion.data()),
                       expansion.order());
    }

    void backward_transform(
        RealZernikeSpan<const std::array<double, 2>, zernike_norm_param, sh_norm_param, sh_phase_param> expansion,
        BallGLQGridSpan<double, GridLayout> values) {
        m_plan.backward(reinterpret_cast<const double*>(expansion.data()), expansion.order(), values.data(),
                        values.order());
    }

    [[nodiscard]] RealZernikeExpansion<zernike_norm_param, sh_norm_param, sh_phase_param> forward_transform(
        BallGLQGridSpan<const double, GridLayout> values, std::size_t order) {
        RealZernikeExpansion<zernike_norm_param, sh_norm_param, sh_phase_param> expansion(order);
        forward_transform(values, expansion);
        return expansion;
    }

    [[nodiscard]] BallGLQGrid<double, GridLayout> backward_transform(
        RealZernikeSpan<const std::array<double, 2>, zernike_norm_param, sh_norm_param, sh_phase_param> expansion,
        std::size_t order) {
        BallGLQGrid<double, GridLayout> grid(order);
        backward_transform(expansion, grid);
        return grid;
    }

    // The underlying plan, for batched device-memory calls through include/zgl.h.
    [[nodiscard]] zgl_plan* plan() const noexcept { return m_plan.get(); }

private:
    static zest_amd::Plan make_plan(std::size_t order) {
        return zest_amd::Plan(ZGL_ZERNIKE, zernike_norm_param == ZernikeNorm::normed ? ZGL_ZNORM_NORMED : ZGL_ZNORM_UNNORMED,
                              sh_norm_param == st::SHNorm::qm ? ZGL_SHNORM_QM : ZGL_SHNORM_GEO,
                              static_cast<int>(sh_phase_param), 0, order);
    }
    zest_amd::Plan m_plan;
};

template <typename GridLayout = DefaultLayout>
using GLQTransformerAcoustics = GLQTransformer<ZernikeNorm::unnormed, st::SHNorm::qm, st::SHPhase::none, GridLayout>;
template <typename GridLayout = DefaultLayout>
using GLQTransformerNormalAcoustics = GLQTransformer<ZernikeNorm::normed, st::SHNorm::qm, st::SHPhase::none, GridLayout>;
template <typename GridLayout = DefaultLayout>
using GLQTransformerQM = GLQTransformer<ZernikeNorm::unnormed, st::SHNorm::qm, st::SHPhase::cs, GridLayout>;
template <typename GridLayout = DefaultLayout>
using GLQTransformerNormalQM = GLQTransformer<ZernikeNorm::normed, st::SHNorm::qm, st::SHPhase::cs, GridLayout>;
template <typename GridLayout = DefaultLayout>
using GLQTransformerGeo = GLQTransformer<ZernikeNorm::unnormed, st::SHNorm::geo, st::SHPhase::none, GridLayout>;
template <typename GridLayout = DefaultLayout>
using GLQTransformerNormalGeo = GLQTransformer<ZernikeNorm::normed, st::SHNorm::geo, st::SHPhase::none, GridLayout>;

template <typename Func>
concept cartesian_function = requires(Func f, std::array<double, 3> x) {
    { f(x) } -> std::same_as<double>;
};

template <typename Func>
concept spherical_function = requires(Func f, double lon, double colat, double r) {
    { f(lon, colat, r) } -> std::same_as<double>;
};

// Expansion of a function on a ball of any radius (zest: 1053-1178).
template <ZernikeNorm zernike_norm_param, st::SHNorm sh_norm_param, st::SHPhase sh_phase_param,
          typename GridLayoutType = DefaultLayout>
class ZernikeTransformer {
public:
    using GridLayout = GridLayoutType;
    using Expansion = RealZernikeExpansion<zernike_norm_param, sh_norm_param, sh_phase_param>;
    using ExpansionSpan = RealZernikeSpan<std::array<double, 2>, zernike_norm_param, sh_norm_param, sh_phase_param>;

    ZernikeTransformer() = default;
    explicit ZernikeTransformer(std::size_t order) : m_grid(order), m_points(order), m_transformer(order) {}

    void resize(std::size_t order) {
        m_points.resize(order);
        m_grid.resize(order);
        m_transformer.resize(order);
    }

    template <spherical_function FuncType>
    void transform(FuncType&& f, double radius, ExpansionSpan expansion) {
        resize(expansion.order());
        m_points.generate_values(m_grid, scaled(f, radius));
        m_transformer.forward_transform(m_grid, expansion);
    }
    template <spherical_function FuncType>
    [[nodiscard]] Expansion transform(FuncType&& f, double radius, std::size_t order) {
        resize(order);
        m_points.generate_values(m_grid, scaled(f, radius));
        return m_transformer.forward_transform(m_grid, order);
    }
    template <cartesian_function FuncType>
    void transform(FuncType&& f, double radius, ExpansionSpan expansion) {
        resize(expansion.order());
        m_points.generate_values(m_grid, cartesian(f, radius));
        m_transformer.forward_transform(m_grid, expansion);
    }
    template <cartesian_function FuncType>
    [[nodiscard]] Expansion transform(FuncType&& f, double radius, std::size_t order) {
        resize(order);
        m_points.generate_values(m_grid, cartesian(f, radius));
        return m_transformer.forward_transform(m_grid, order);
    }

private:
    template <typename F>
    static auto scaled(F& f, double radius) {
        return [&f, radius](double lon, double colat, double r) { return f(lon, colat, r * radius); };
    }
    template <typename F>
    static auto cartesian(F& f, double radius) {
        return [&f, radius](double lon, double colat, double r) {
            const double rad = r * radius;
            const double s = std::sin(colat);
            const std::array<double, 3> x = {rad * s * std::cos(lon), rad * s * std::sin(lon), rad * std::cos(colat)};
            return f(x);
        };
    }

    BallGLQGrid<double, GridLayout> m_grid;
    BallGLQGridPoints<GridLayout> m_points;
    GLQTransformer<zernike_norm_param, sh_norm_param, sh_phase_param, GridLayout> m_transformer;
};

template <typename GridLayout = DefaultLayout>
using ZernikeTransformerAcoustics = ZernikeTransformer<ZernikeNorm::unnormed, st::SHNorm::qm, st::SHPhase::none, GridLayout>;
template <typename GridLayout = DefaultLayout>
using ZernikeTransformerNormalAcoustics = ZernikeTransformer<ZernikeNorm::normed, st::SHNorm::qm, st::SHPhase::none, GridLayout>;
template <typename GridLayout = DefaultLayout>
using ZernikeTransformerQM = ZernikeTransformer<ZernikeNorm::unnormed, st::SHNorm::qm, st::SHPhase::cs, GridLayout>;
template <typename GridLayout = DefaultLayout>
using ZernikeTransformerNormalQM = ZernikeTransformer<ZernikeNorm::normed, st::SHNorm::qm, st::SHPhase::cs, GridLayout>;
template <typename GridLayout = DefaultLayout>
using ZernikeTransformerGeo = ZernikeTransformer<ZernikeNorm::unnormed, st::SHNorm::geo, st::SHPhase::none, GridLayout>;
template <typename GridLayout = DefaultLayout>
using ZernikeTransformerNormalGeo = ZernikeTransformer<ZernikeNorm::normed, st::SHNorm::geo, st::SHPhase::none, GridLayout>;

}  // namespace zt
}  // namespace zest
