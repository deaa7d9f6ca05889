// Drop-in replacement of zest's "rotor.hpp" (sebsassi/zest, include/rotor.hpp)
// backed by the MI355X library (include/zgl.h). RotationType,
// euler_angles_from_rotation_matrix and Rotor keep zest's names and
// signatures; Rotor::rotate and Rotor::polar_rotate run zgl_rotate /
// zgl_polar_rotate on the caller's host coefficients (staged through the GPU,
// synchronous, in place, as zest works in place). The WignerdPiHalfCollection
// argument is accepted for compatibility; the library keeps its own copy of
// the same table on the device.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <memory>
#include <numbers>
#include <type_traits>
#include <vector>

#include "real_sh_expansion.hpp"
#include "zernike_expansion.hpp"
#include "wignerd_collection.hpp"

#include "zgl_plan.hpp"

namespace zest {

/* Whether a rotation applies to the object or the coordinate system
   (zest: include/rotor.hpp:37-43). */
enum class RotationType { object, coordinate };

/* Euler angles (alpha, beta, gamma) of a rotation matrix
   (zest: include/rotor.hpp:49-60, computed the same way). */
[[nodiscard]] inline std::array<double, 3> euler_angles_from_rotation_matrix(
    const std::array<std::array<double, 3>, 3>& rot) noexcept {
    std::array<double, 3> a{};
    const double m[9] = {rot[0][0], rot[0][1], rot[0][2], rot[1][0], rot[1][1],
                         rot[1][2], rot[2][0], rot[2][1], rot[2][2]};
    zgl_euler_angles_from_rotation_matrix(m, a.data());
    return a;
}

/*
Rotations of spherical harmonic and Zernike coefficients
(zest: include/rotor.hpp:91-403). Intrinsic ZYZ Euler angles.
*/
class Rotor {
public:
    Rotor() = default;
    explicit Rotor(std::size_t max_order) : m_max_order(max_order) {}
    Rotor(const Rotor& o) : m_max_order(o.m_max_order) {}
    Rotor& operator=(const Rotor& o) {
        m_max_order = o.m_max_order;
        return *this;
    }
    Rotor(Rotor&&) noexcept = default;
    Rotor& operator=(Rotor&&) noexcept = default;

    void expand(std::size_t max_order) { m_max_order = std::max(m_max_order, max_order); }
    [[nodiscard]] std::size_t max_order() const noexcept { return m_max_order; }

    template <st::real_sh_expansion ExpansionType>
        requires std::same_as<typename std::remove_cvref_t<ExpansionType>::value_type, std::array<double, 2>>
    void rotate(ExpansionType&& expansion, [[maybe_unused]] const WignerdPiHalfCollection& wigner_d_pi2,
                const std::array<double, 3>& euler_angles, RotationType type) {
        expand(expansion.order());
        run(ZGL_SH, data_of(expansion), expansion.order(), euler_angles.data(), type, false);
    }

    // An SH span over every other degree (the (l, m) block of one Zernike
    // degree n = order - 1): rotated as that block of a Zernike expansion.
    template <st::row_skipping_real_sh_expansion ExpansionType>
        requires std::same_as<typename std::remove_cvref_t<ExpansionType>::value_type, std::array<double, 2>>
    void rotate(ExpansionType&& expansion, [[maybe_unused]] const WignerdPiHalfCollection& wigner_d_pi2,
                const std::array<double, 3>& euler_angles, RotationType type) {
        row_skipping(expansion, euler_angles.data(), type, false);
    }

    template <zt::real_zernike_expansion ExpansionType>
        requires std::same_as<typename std::remove_cvref_t<ExpansionType>::value_type, std::array<double, 2>>
    void rotate(ExpansionType&& expansion, [[maybe_unused]] const WignerdPiHalfCollection& wigner_d_pi2,
                const std::array<double, 3>& euler_angles, RotationType type) {
        expand(expansion.order());
        run(ZGL_ZERNIKE, data_of(expansion), expansion.order(), euler_angles.data(), type, false);
    }

    template <st::real_sh_expansion ExpansionType>
        requires std::same_as<typename std::remove_cvref_t<ExpansionType>::value_type, std::array<double, 2>>
    void rotate(ExpansionType&& expansion, const WignerdPiHalfCollection& wigner_d_pi2,
                const std::array<std::array<double, 3>, 3>& matrix, RotationType type) {
        rotate(expansion, wigner_d_pi2, euler_angles_from_rotation_matrix(matrix), type);
    }

    template <zt::real_zernike_expansion ExpansionType>
        requires std::same_as<typename std::remove_cvref_t<ExpansionType>::value_type, std::array<double, 2>>
    void rotate(ExpansionType&& expansion, const WignerdPiHalfCollection& wigner_d_pi2,
                const std::array<std::array<double, 3>, 3>& matrix, RotationType type) {
        rotate(expansion, wigner_d_pi2, euler_angles_from_rotation_matrix(matrix), type);
    }

    template <st::real_sh_expansion ExpansionType>
        requires std::same_as<typename std::remove_cvref_t<ExpansionType>::value_type, std::array<double, 2>>
    void polar_rotate(ExpansionType&& expansion, double angle, RotationType type) {
        expand(expansion.order());
        run(ZGL_SH, data_of(expansion), expansion.order(), &angle, type, true);
    }

    template <zt::real_zernike_expansion ExpansionType>
        requires std::same_as<typename std::remove_cvref_t<ExpansionType>::value_type, std::array<double, 2>>
    void polar_rotate(ExpansionType&& expansion, double angle, RotationType type) {
        expand(expansion.order());
        run(ZGL_ZERNIKE, data_of(expansion), expansion.order(), &angle, type, true);
    }

private:
    template <class E>
    static double* data_of(E& expansion) {
        if constexpr (requires { expansion.flatten(); })
            return reinterpret_cast<double*>(expansion.flatten().data());
        else
            return reinterpret_cast<double*>(expansion.data());
    }

    zgl_plan* plan_for(zgl_kind kind) {
        auto& p = kind == ZGL_ZERNIKE ? m_zt : m_st;
        if (!p)
            p = std::make_unique<zest_amd::Plan>(kind, ZGL_ZNORM_NORMED, ZGL_SHNORM_GEO, ZGL_PHASE_NONE,
                                                 ZGL_LAYOUT_LONLAT, 0);
        return p->get();
    }

    void run(zgl_kind kind, double* coeffs, std::size_t order, const double* angles, RotationType type,
             bool polar) {
        const int t = type == RotationType::object ? 0 : 1;
        zgl_plan* p = plan_for(kind);
        if (polar)
            zest_amd::zgl_check(zgl_polar_rotate(p, coeffs, order, angles, t, 1, 0, ZGL_MEM_HOST, nullptr));
        else
            zest_amd::zgl_check(zgl_rotate(p, coeffs, order, angles, t, 1, 0, ZGL_MEM_HOST, nullptr));
    }

    template <class E>
    void row_skipping(E& expansion, const double* angles, RotationType type, bool polar) {
        const std::size_t order = expansion.order();
        if (order == 0) return;
        expand(order);
        // degree n = order - 1 occupies pairs [tet(n, n & 1, 0), tet(order, order & 1, 0))
        const std::size_t n = order - 1;
        const std::size_t first = (n + 1) * (n + 3) * (2 * n + 1) / 24;
        const std::size_t count = ((order + 1) * (order + 1)) >> 2;
        std::vector<double> full(2 * (first + count), 0.0);
        double* block = data_of(expansion);
        std::copy(block, block + 2 * count, full.begin() + 2 * first);
        run(ZGL_ZERNIKE, full.data(), order, angles, type, polar);
        std::copy(full.begin() + 2 * first, full.begin() + 2 * (first + count), block);
    }

    std::size_t m_max_order{};
    std::unique_ptr<zest_amd::Plan> m_st;
    std::unique_ptr<zest_amd::Plan> m_zt;
};

}  // namespace zest
