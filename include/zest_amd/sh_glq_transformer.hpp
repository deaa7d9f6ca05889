// Drop-in replacement of zest's "sh_glq_transformer.hpp" (sebsassi/zest,
// include/sh_glq_transformer.hpp) backed by the MI355X library
// (include/zgl.h). Names, namespaces, template parameters, layouts and
// convention types are zest's; zest's data-model headers are used as they
// are. GLQTransformer::forward_transform / backward_transform (zest: 632-754)
// call zgl_forward / zgl_backward on the GPU and return when the output is
// written.
//
// Supported grid layouts: LonLatLayout<CacheLineAlignment> (zest's default)
// and LatLonLayout<CacheLineAlignment>.
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
#include "real_sh_expansion.hpp"

#include "zgl_plan.hpp"

namespace zest {
namespace st {

// [lat][lon] grid, longitudes padded to a multiple of the vector size
// (zest: include/sh_glq_transformer.hpp:48-118).
template <typename AlignmentType = CacheLineAlignment>
struct LatLonLayout {
    using Alignment = AlignmentType;
    [[nodiscard]] static constexpr std::size_t lat_size(std::size_t order) noexcept { return order; }
    [[nodiscard]] static constexpr std::size_t lon_size(std::size_t order) noexcept {
        const std::size_t n = order ? 2 * order - 1 : 0;
        if constexpr (std::is_same_v<Alignment, NoAlignment>)
            return n;
        else
            return zest::detail::next_divisible<Alignment::template vector_size<double>()>(n);
    }
    [[nodiscard]] static constexpr std::size_t size(std::size_t order) noexcept {
        return lat_size(order) * lon_size(order);
    }
    [[nodiscard]] static constexpr std::array<std::size_t, 2> shape(std::size_t order) noexcept {
        return {lat_size(order), lon_size(order)};
    }
    [[nodiscard]] static constexpr std::size_t fft_size(std::size_t order) noexcept {
        return (lon_size(order) >> 1) + 1;
    }
    [[nodiscard]] static constexpr std::array<std::size_t, 2> fft_stride(std::size_t order) noexcept {
        return {fft_size(order), 1};
    }
    static constexpr std::size_t lat_axis = 0;
    static constexpr std::size_t lon_axis = 1;
};

// [lon][lat] grid, latitudes padded to a multiple of the vector size
// (zest: include/sh_glq_transformer.hpp:125-195).
template <typename AlignmentType = CacheLineAlignment>
struct LonLatLayout {
    using Alignment = AlignmentType;
    [[nodiscard]] static constexpr std::size_t lat_size(std::size_t order) noexcept {
        if constexpr (std::is_same_v<Alignment, NoAlignment>)
            return order;
        else
            return zest::detail::next_divisible<Alignment::template vector_size<double>()>(order);
    }
    [[nodiscard]] static constexpr std::size_t lon_size(std::size_t order) noexcept {
        return order ? 2 * order - 1 : 0;
    }
    [[nodiscard]] static constexpr std::size_t size(std::size_t order) noexcept {
        return lat_size(order) * lon_size(order);
    }
    [[nodiscard]] static constexpr std::array<std::size_t, 2> shape(std::size_t order) noexcept {
        return {lon_size(order), lat_size(order)};
    }
    [[nodiscard]] static constexpr std::size_t fft_size(std::size_t order) noexcept {
        return (lon_size(order) >> 1) + 1;
    }
    [[nodiscard]] static constexpr std::array<std::size_t, 2> fft_stride(std::size_t order) {
        return {lat_size(order), 1};
    }
    static constexpr std::size_t lat_axis = 1;
    static constexpr std::size_t lon_axis = 0;
};

using DefaultLayout = LonLatLayout<>;

template <typename ElementType, typename LayoutType = DefaultLayout>
class SphereGLQGridSpan : public MDSpan<ElementType, 2> {
public:
    using Base = MDSpan<ElementType, 2>;
    using typename Base::element_type;
    using Layout = LayoutType;
    using ConstView = SphereGLQGridSpan<const element_type, Layout>;
    using Base::data;
    using Base::extents;
    using Base::size;

    [[nodiscard]] static constexpr std::size_t size(std::size_t order) noexcept { return Layout::size(order); }
    [[nodiscard]] static constexpr std::array<std::size_t, 2> shape(std::size_t order) noexcept {
        return Layout::shape(order);
    }

    constexpr SphereGLQGridSpan() noexcept = default;
    constexpr SphereGLQGridSpan(element_type* data, std::size_t order) noexcept
        : Base(data, Layout::shape(order)), m_order(order) {}
    constexpr SphereGLQGridSpan(std::span<element_type> buffer, std::size_t order) noexcept
        : Base(buffer.data(), Layout::shape(order)), m_order(order) {}

    [[nodiscard]] constexpr std::size_t order() const noexcept { return m_order; }
    [[nodiscard]] constexpr const std::array<std::size_t, 2>& shape() const noexcept { return extents(); }
    [[nodiscard]] constexpr std::span<element_type> flatten() const noexcept {
        return std::span<element_type>(data(), size());
    }
    // Through MDSpan's (data, size, extents) constructor, which stays usable
    // in constant expressions.
    [[nodiscard]] constexpr operator ConstView() const noexcept {
        return ConstView(data(), size(), extents(), m_order);
    }

private:
    friend SphereGLQGridSpan<std::remove_const_t<element_type>, Layout>;
    constexpr SphereGLQGridSpan(element_type* data, std::size_t size, const std::array<std::size_t, 2>& extents,
                    std::size_t order) noexcept
        : Base(data, size, extents), m_order(order) {}

    std::size_t m_order{};
};

template <typename ElementType, typename LayoutType = DefaultLayout>
class SphereGLQGrid {
public:
    using element_type = ElementType;
    using value_type = std::remove_cvref_t<element_type>;
    using Layout = LayoutType;
    using View = SphereGLQGridSpan<element_type, Layout>;
    using ConstView = SphereGLQGridSpan<const element_type, Layout>;

    SphereGLQGrid() = default;
    explicit SphereGLQGrid(std::size_t order)
        : m_values(Layout::size(order)), m_shape(Layout::shape(order)), m_order(order) {}

    [[nodiscard]] std::size_t order() const noexcept { return m_order; }
    [[nodiscard]] std::array<std::size_t, 2> shape() const noexcept { return m_shape; }
    [[nodiscard]] std::span<const element_type> flatten() const noexcept { return m_values; }
    std::span<element_type> flatten() noexcept { return m_values; }

    [[nodiscard]] operator View() noexcept { return View(m_values, m_order); }
    [[nodiscard]] operator ConstView() const noexcept { return ConstView(m_values, m_order); }

    void resize(std::size_t order) {
        m_values.resize(Layout::size(order));
        m_shape = Layout::shape(order);
        m_order = order;
    }

    [[nodiscard]] element_type operator()(std::size_t i, std::size_t j) const noexcept {
        return m_values[i * m_shape[1] + j];
    }
    [[nodiscard]] element_type& operator()(std::size_t i, std::size_t j) noexcept {
        return m_values[i * m_shape[1] + j];
    }

private:
    std::vector<element_type, AlignedAllocator<element_type, typename LayoutType::Alignment>> m_values{};
    std::array<std::size_t, 2> m_shape{};
    std::size_t m_order{};
};

template <typename T>
concept sphere_glq_grid =
    std::same_as<std::remove_cvref_t<T>, SphereGLQGridSpan<typename std::remove_cvref_t<T>::element_type,
                                                            typename std::remove_cvref_t<T>::Layout>> ||
    std::same_as<std::remove_cvref_t<T>, SphereGLQGrid<typename std::remove_cvref_t<T>::element_type,
                                                        typename std::remove_cvref_t<T>::Layout>>;

// Grid sample points and function sampling (zest: include/sh_glq_transformer.hpp:384-504).
template <typename LayoutType = DefaultLayout>
class SphereGLQGridPoints {
public:
    using GridLayout = LayoutType;

    SphereGLQGridPoints() = default;
    explicit SphereGLQGridPoints(std::size_t order) { resize(order); }

    void resize(std::size_t order) {
        const auto s = GridLayout::shape(order);
        resize(s[GridLayout::lon_axis], s[GridLayout::lat_axis]);
    }

    [[nodiscard]] std::array<std::size_t, 2> shape() noexcept { return {m_nodes.size(), m_lon.size()}; }
    [[nodiscard]] std::span<const double> longitudes() const noexcept { return m_lon; }
    [[nodiscard]] std::span<const double> glq_nodes() const noexcept { return m_nodes; }

    template <sphere_glq_grid GridType, typename FuncType>
        requires std::same_as<typename std::remove_cvref_t<GridType>::Layout, GridLayout>
    void generate_values(GridType&& grid, FuncType&& f) {
        resize(grid.order());
        if constexpr (GridLayout::lat_axis == 0) {
            for (std::size_t i = 0; i < m_nodes.size(); ++i)
                for (std::size_t j = 0; j < m_lon.size(); ++j) grid(i, j) = f(m_lon[j], m_nodes[i]);
        } else {
            for (std::size_t i = 0; i < m_lon.size(); ++i)
                for (std::size_t j = 0; j < m_nodes.size(); ++j) grid(i, j) = f(m_lon[i], m_nodes[j]);
        }
    }

    template <typename FuncType>
    auto generate_values(FuncType&& f, std::size_t order) {
        using CodomainType = std::invoke_result_t<FuncType, double, double>;
        SphereGLQGrid<CodomainType, GridLayout> grid(order);
        generate_values(grid, f);
        return grid;
    }

private:
    void resize(std::size_t nlon, std::size_t nlat) {
        if (nlon != m_lon.size()) {
            m_lon.resize(nlon);
            const double dlon = (2.0 * std::numbers::pi) / double(nlon);
            for (std::size_t i = 0; i < nlon; ++i) m_lon[i] = dlon * double(i);
        }
        if (nlat != m_nodes.size()) {
            m_nodes.resize(nlat);
            gl::gl_nodes<gl::UnpackedLayout, gl::GLNodeStyle::angle>(m_nodes, nlat & 1);
        }
    }

    std::vector<double> m_lon;
    std::vector<double> m_nodes;
};

// Sphere grid <-> real spherical-harmonic expansion (zest: 513-1213), on the GPU.
template <SHNorm sh_norm_param, SHPhase sh_phase_param, typename GridLayoutType = DefaultLayout>
class GLQTransformer {
    static constexpr bool lonlat = std::same_as<GridLayoutType, LonLatLayout<CacheLineAlignment>>;
    static_assert(lonlat || std::same_as<GridLayoutType, LatLonLayout<CacheLineAlignment>>,
                  "the MI355X transformer supports LonLatLayout<> and LatLonLayout<>");

public:
    using GridLayout = GridLayoutType;
    using SHLayout = TriangleLayout<IndexingMode::nonnegative>;
    static constexpr SHNorm norm = sh_norm_param;
    static constexpr SHPhase phase = sh_phase_param;

    GLQTransformer() : m_plan(make_plan(0)) {}
    explicit GLQTransformer(std::size_t order) : m_plan(make_plan(order)) {}

    [[nodiscard]] std::size_t order() const noexcept { return m_plan.order(); }
    void resize(std::size_t order) { m_plan.resize(order); }

    void forward_transform(SphereGLQGridSpan<const double, GridLayout> values,
                           RealSHSpan<std::array<double, 2>, sh_norm_param, sh_phase_param> expansion) {
        m_plan.forward(values.data(), values.order(), reinterpret_cast<double*>(expansion.data()),
                       expansion.order());
    }

    void backward_transform(RealSHSpan<const std::array<double, 2>, sh_norm_param, sh_phase_param> expansion,
                            SphereGLQGridSpan<double, GridLayout> values) {
        m_plan.backward(reinterpret_cast<const double*>(expansion.data()), expansion.order(), values.data(),
                        values.order());
    }

    // Backward transform of an expansion holding only even or only odd l
    // (zest: 680-700, 1053-1133): the missing rows are zero, so it equals the
    // backward transform of the full expansion with those rows zeroed. As in
    // zest, the rows read are those of a view of the same data truncated to
    // min(expansion order, grid order): that view's order fixes the parity,
    // and row l sits at (l*l)/4 in the row-skipping layout whichever parity
    // it has.
    template <row_skipping_real_sh_expansion Expansion>
        requires(std::remove_cvref_t<Expansion>::norm == sh_norm_param) &&
                (std::remove_cvref_t<Expansion>::phase == sh_phase_param) &&
                std::same_as<typename std::remove_cvref_t<Expansion>::value_type, std::array<double, 2>> &&
                has_parity<Expansion>
    void backward_transform(Expansion&& expansion, SphereGLQGridSpan<double, GridLayout> values) {
        const std::size_t order = std::min<std::size_t>(expansion.order(), values.order());
        typename std::remove_cvref_t<Expansion>::ConstView truncated(expansion.data(), order);
        RealSHExpansion<sh_norm_param, sh_phase_param> full(order);
        for (auto l : truncated.indices()) {
            auto row = truncated[l];
            for (std::size_t m = 0; m <= std::size_t(l); ++m) full(l, m) = row[m];
        }
        backward_transform(full, values);
    }

    [[nodiscard]] RealSHExpansion<sh_norm_param, sh_phase_param> forward_transform(
        SphereGLQGridSpan<const double, GridLayout> values, std::size_t order) {
        RealSHExpansion<sh_norm_param, sh_phase_param> expansion(order);
        forward_transform(values, expansion);
        return expansion;
    }

    [[nodiscard]] SphereGLQGrid<double, GridLayout> backward_transform(
        RealSHSpan<const std::array<double, 2>, sh_norm_param, sh_phase_param> expansion, std::size_t order) {
        SphereGLQGrid<double, GridLayout> grid(order);
        backward_transform(expansion, grid);
        return grid;
    }

    template <row_skipping_real_sh_expansion Expansion>
        requires(std::remove_cvref_t<Expansion>::norm == sh_norm_param) &&
                (std::remove_cvref_t<Expansion>::phase == sh_phase_param) &&
                std::same_as<typename std::remove_cvref_t<Expansion>::value_type, std::array<double, 2>> &&
                has_parity<Expansion>
    [[nodiscard]] SphereGLQGrid<double, GridLayout> backward_transform(Expansion&& expansion, std::size_t order) {
        SphereGLQGrid<double, GridLayout> grid(order);
        backward_transform(std::forward<Expansion>(expansion), grid);
        return grid;
    }

    // The underlying plan, for batched device-memory calls through include/zgl.h.
    [[nodiscard]] zgl_plan* plan() const noexcept { return m_plan.get(); }

private:
    static zest_amd::Plan make_plan(std::size_t order) {
        return zest_amd::Plan(ZGL_SH, ZGL_ZNORM_UNNORMED, sh_norm_param == SHNorm::qm ? ZGL_SHNORM_QM : ZGL_SHNORM_GEO,
                              static_cast<int>(sh_phase_param), lonlat ? ZGL_LAYOUT_LONLAT : ZGL_LAYOUT_LATLON,
                              order);
    }
    zest_amd::Plan m_plan;
};

template <typename GridLayout = DefaultLayout>
using GLQTransformerAcoustics = GLQTransformer<SHNorm::qm, SHPhase::none, GridLayout>;
template <typename GridLayout = DefaultLayout>
using GLQTransformerQM = GLQTransformer<SHNorm::qm, SHPhase::cs, GridLayout>;
template <typename GridLayout = DefaultLayout>
using GLQTransformerGeo = GLQTransformer<SHNorm::geo, SHPhase::none, GridLayout>;

template <typename Func>
concept cartesian_function = requires(Func f, std::array<double, 3> x) {
    { f(x) } -> std::same_as<double>;
};

template <typename Func>
concept spherical_function = requires(Func f, double lon, double colat) {
    { f(lon, colat) } -> std::same_as<double>;
};

// Expansion of a function on the unit sphere (zest: 1267-1379).
template <SHNorm sh_norm_param, SHPhase sh_phase_param, typename GridLayoutType = DefaultLayout>
class SHTransformer {
public:
    using GridLayout = GridLayoutType;
    using Expansion = RealSHExpansion<sh_norm_param, sh_phase_param>;
    using ExpansionSpan = RealSHSpan<std::array<double, 2>, sh_norm_param, sh_phase_param>;

    SHTransformer() = default;
    explicit SHTransformer(std::size_t order) : m_grid(order), m_points(order), m_transformer(order) {}

    void resize(std::size_t order) {
        m_points.resize(order);
        m_grid.resize(order);
        m_transformer.resize(order);
    }

    template <spherical_function FuncType>
    void transform(FuncType&& f, ExpansionSpan expansion) {
        resize(expansion.order());
        m_points.generate_values(m_grid, f);
        m_transformer.forward_transform(m_grid, expansion);
    }
    template <spherical_function FuncType>
    [[nodiscard]] Expansion transform(FuncType&& f, std::size_t order) {
        resize(order);
        m_points.generate_values(m_grid, f);
        return m_transformer.forward_transform(m_grid, order);
    }
    template <cartesian_function FuncType>
    void transform(FuncType&& f, ExpansionSpan expansion) {
        resize(expansion.order());
        m_points.generate_values(m_grid, spherical(f));
        m_transformer.forward_transform(m_grid, expansion);
    }
    template <cartesian_function FuncType>
    [[nodiscard]] Expansion transform(FuncType&& f, std::size_t order) {
        resize(order);
        m_points.generate_values(m_grid, spherical(f));
        return m_transformer.forward_transform(m_grid, order);
    }

private:
    template <typename F>
    static auto spherical(F& f) {
        return [&f](double lon, double colat) {
            const double s = std::sin(colat);
            const std::array<double, 3> x = {s * std::cos(lon), s * std::sin(lon), std::cos(colat)};
            return f(x);
        };
    }

    SphereGLQGrid<double, GridLayout> m_grid;
    SphereGLQGridPoints<GridLayout> m_points;
    GLQTransformer<sh_norm_param, sh_phase_param, GridLayout> m_transformer;
};

template <typename GridLayout = DefaultLayout>
using SHTransformerAcoustics = SHTransformer<SHNorm::qm, SHPhase::none, GridLayout>;
template <typename GridLayout = DefaultLayout>
using SHTransformerQM = SHTransformer<SHNorm::qm, SHPhase::cs, GridLayout>;
template <typename GridLayout = DefaultLayout>
using SHTransformerGeo = SHTransformer<SHNorm::geo, SHPhase::none, GridLayout>;

}  // namespace st
}  // namespace zest
