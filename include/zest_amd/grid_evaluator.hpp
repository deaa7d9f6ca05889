// Drop-in replacement of zest's "grid_evaluator.hpp" (sebsassi/zest,
// include/grid_evaluator.hpp) backed by the MI355X library (include/zgl.h).
// Put this directory ahead of zest's include directory: the class names,
// constructors, resize() and evaluate() signatures are zest's, and the
// expansion types are zest's own (real_sh_expansion.hpp, zernike_expansion.hpp).
//
// evaluate() calls zgl_evaluate on host memory: the coefficients and points
// are staged through the GPU and the values returned in a std::vector, shaped
// [lon][lat] (SH) or [lon][lat][rad] (Zernike) in row-major order, as zest
// returns them. The expansion's conventions select the library plan, which
// is created on first use and again when the conventions change.
#pragma once

#include <cstddef>
#include <memory>
#include <span>
#include <type_traits>
#include <vector>

#include "real_sh_expansion.hpp"
#include "zernike_expansion.hpp"

#include "zgl_plan.hpp"

namespace zest {

namespace detail_zgl {

template <class E>
const double* coeff_pointer(const E& expansion) {
    if constexpr (requires { expansion.flatten(); })
        return reinterpret_cast<const double*>(expansion.flatten().data());
    else
        return reinterpret_cast<const double*>(expansion.data());
}

// One plan per evaluator, rebuilt when the requested conventions change.
class EvalPlan {
public:
    zgl_plan* get(zgl_kind kind, int zn, int sn, int sp) {
        if (!m_plan || kind != m_kind || zn != m_zn || sn != m_sn || sp != m_sp) {
            m_plan = std::make_unique<zest_amd::Plan>(kind, zn, sn, sp, ZGL_LAYOUT_LONLAT, 0);
            m_kind = kind;
            m_zn = zn;
            m_sn = sn;
            m_sp = sp;
        }
        return m_plan->get();
    }
    EvalPlan() = default;
    EvalPlan(const EvalPlan&) {}  // copies create their own plan on first use
    EvalPlan& operator=(const EvalPlan&) { return *this; }
    EvalPlan(EvalPlan&&) noexcept = default;
    EvalPlan& operator=(EvalPlan&&) noexcept = default;

private:
    std::unique_ptr<zest_amd::Plan> m_plan;
    zgl_kind m_kind = ZGL_SH;
    int m_zn = 0, m_sn = 0, m_sp = 0;
};

}  // namespace detail_zgl

namespace st {

/*
Class for evaluating spherical harmonic expansions on arbitrary grids
(zest: include/grid_evaluator.hpp:57-173).
*/
class GridEvaluator {
public:
    GridEvaluator() = default;
    explicit GridEvaluator(std::size_t max_order) : m_max_order(max_order) {}
    GridEvaluator(std::size_t max_order, std::size_t lon_size, std::size_t lat_size)
        : m_lon_size(lon_size), m_lat_size(lat_size), m_max_order(max_order) {}

    // The library keeps its buffers in the plan and grows them on demand;
    // resize only records the sizes, as the reference's accessors report them.
    void resize(std::size_t max_order, std::size_t lon_size, std::size_t lat_size) {
        m_max_order = max_order;
        m_lon_size = lon_size;
        m_lat_size = lat_size;
    }

    template <real_sh_expansion ExpansionType>
    [[nodiscard]] std::vector<double> evaluate(ExpansionType&& expansion,
                                               std::span<const double> longitudes,
                                               std::span<const double> colatitudes) {
        using E = std::remove_cvref_t<ExpansionType>;
        if (longitudes.size() == 0 || colatitudes.size() == 0) return std::vector<double>{};
        const std::size_t order = expansion.order();
        resize(order, longitudes.size(), colatitudes.size());
        zgl_plan* plan = m_plan.get(ZGL_SH, ZGL_ZNORM_NORMED,
                                    E::norm == SHNorm::qm ? ZGL_SHNORM_QM : ZGL_SHNORM_GEO,
                                    E::phase == SHPhase::cs ? ZGL_PHASE_CS : ZGL_PHASE_NONE);
        std::vector<double> res(longitudes.size() * colatitudes.size());
        zest_amd::zgl_check(zgl_evaluate(plan, detail_zgl::coeff_pointer(expansion), order,
                                         longitudes.data(), longitudes.size(), colatitudes.data(),
                                         colatitudes.size(), nullptr, 1, res.data(), 1, 0, 0,
                                         ZGL_MEM_HOST, nullptr));
        return res;
    }

private:
    detail_zgl::EvalPlan m_plan;
    std::size_t m_lon_size{};
    std::size_t m_lat_size{};
    std::size_t m_max_order{};
};

}  // namespace st

namespace zt {

/*
Class for evaluating Zernike expansions on arbitrary grids
(zest: include/grid_evaluator.hpp:175-324).
*/
class GridEvaluator {
public:
    GridEvaluator() = default;
    explicit GridEvaluator(std::size_t max_order) : m_max_order(max_order) {}
    GridEvaluator(std::size_t max_order, std::size_t lon_size, std::size_t lat_size,
                  std::size_t rad_size)
        : m_lon_size(lon_size), m_lat_size(lat_size), m_rad_size(rad_size), m_max_order(max_order) {}

    void resize(std::size_t max_order, std::size_t lon_size, std::size_t lat_size,
                std::size_t rad_size) {
        m_max_order = max_order;
        m_lon_size = lon_size;
        m_lat_size = lat_size;
        m_rad_size = rad_size;
    }

    template <real_zernike_expansion ExpansionType>
    [[nodiscard]] std::vector<double> evaluate(ExpansionType&& expansion,
                                               std::span<const double> longitudes,
                                               std::span<const double> colatitudes,
                                               std::span<const double> radii) {
        using E = std::remove_cvref_t<ExpansionType>;
        if (longitudes.size() == 0 || colatitudes.size() == 0 || radii.size() == 0)
            return std::vector<double>{};
        const std::size_t order = expansion.order();
        resize(order, longitudes.size(), colatitudes.size(), radii.size());
        zgl_plan* plan = m_plan.get(
            ZGL_ZERNIKE, E::zernike_norm == ZernikeNorm::normed ? ZGL_ZNORM_NORMED : ZGL_ZNORM_UNNORMED,
            E::sh_norm == st::SHNorm::qm ? ZGL_SHNORM_QM : ZGL_SHNORM_GEO,
            E::sh_phase == st::SHPhase::cs ? ZGL_PHASE_CS : ZGL_PHASE_NONE);
        std::vector<double> res(longitudes.size() * colatitudes.size() * radii.size());
        zest_amd::zgl_check(zgl_evaluate(plan, detail_zgl::coeff_pointer(expansion), order,
                                         longitudes.data(), longitudes.size(), colatitudes.data(),
                                         colatitudes.size(), radii.data(), radii.size(), res.data(), 1,
                                         0, 0, ZGL_MEM_HOST, nullptr));
        return res;
    }

private:
    detail_zgl::EvalPlan m_plan;
    std::size_t m_lon_size{};
    std::size_t m_lat_size{};
    std::size_t m_rad_size{};
    std::size_t m_max_order{};
};

}  // namespace zt

}  // namespace zest
