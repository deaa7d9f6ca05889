// Drop-in replacement of zest's "wignerd_collection.hpp" (sebsassi/zest,
// include/wignerd_collection.hpp) backed by the MI355X library
// (include/zgl.h): WignerdSpan and WignerdPiHalfCollection keep zest's
// names, accessors and table layout; the table is generated on the GPU
// (zgl_wignerd_pi2) and held on the host, as zest holds it. zest's
// per-degree building blocks rotate_l / polar_rotate_l are host functions
// here as in zest, for callers that rotate single degrees themselves; the
// drop-in Rotor (rotor.hpp) runs whole rotations on the GPU.
#pragma once

#include <complex>
#include <cstddef>
#include <span>
#include <vector>

#include "zgl_plan.hpp"

namespace zest {

/*
Non-owning view of a Wigner (small) d-matrix at pi/2
(zest: include/wignerd_collection.hpp:38-66).
*/
template <typename ElementType>
class WignerdSpan {
public:
    constexpr WignerdSpan() noexcept = default;
    constexpr WignerdSpan(ElementType* data, std::size_t order) noexcept
        : m_data(data), m_size(order * order), m_order(order) {}

    [[nodiscard]] constexpr double operator()(std::size_t m1, std::size_t m2) const noexcept {
        return m_data[idx(m1, m2)];
    }

    [[nodiscard]] constexpr std::size_t idx(std::size_t m1, std::size_t m2) const noexcept {
        return m_order * m1 + m2;
    }

private:
    ElementType* m_data{};
    std::size_t m_size{};
    std::size_t m_order{};
};

/*
Collection of Wigner (small) d-matrices at pi/2 of degrees l < max_order
(zest: include/wignerd_collection.hpp:71-132).
*/
class WignerdPiHalfCollection {
public:
    WignerdPiHalfCollection() = default;
    explicit WignerdPiHalfCollection(std::size_t max_order) { expand(max_order); }

    void expand(std::size_t max_order) {
        if (max_order <= m_max_order) return;
        std::vector<double> table((max_order * (max_order + 1) * (2 * max_order + 1)) / 6);
        zest_amd::Plan plan(ZGL_SH, ZGL_ZNORM_NORMED, ZGL_SHNORM_GEO, ZGL_PHASE_NONE, ZGL_LAYOUT_LONLAT, 0);
        zest_amd::zgl_check(zgl_wignerd_pi2(plan.get(), max_order, table.data(), ZGL_MEM_HOST, nullptr));
        m_matrices = std::move(table);
        m_max_order = max_order;
    }

    [[nodiscard]] std::span<const double> matrices() const noexcept { return m_matrices; }
    [[nodiscard]] std::size_t max_order() const noexcept { return m_max_order; }

    [[nodiscard]] double operator()(std::size_t l, std::size_t m1, std::size_t m2) const noexcept {
        return m_matrices[idx(l, m1, m2)];
    }
    [[nodiscard]] WignerdSpan<const double> operator()(std::size_t l) const noexcept {
        return WignerdSpan<const double>(m_matrices.data() + (l * (l + 1) * (2 * l + 1)) / 6, l + 1);
    }
    [[nodiscard]] WignerdSpan<const double> operator[](std::size_t l) const noexcept {
        return WignerdSpan<const double>(m_matrices.data() + (l * (l + 1) * (2 * l + 1)) / 6, l + 1);
    }

private:
    [[nodiscard]] static constexpr std::size_t idx(std::size_t l, std::size_t m1, std::size_t m2) noexcept {
        return (l * (l + 1) * (2 * l + 1)) / 6 + (l + 1) * m1 + m2;
    }

    std::vector<double> m_matrices{};
    std::size_t m_max_order{};
};

namespace detail {

// a * b written out: the operands are finite, so the Inf/NaN recovery of
// std::complex's operator* is not wanted (zest multiplies the same way).
[[nodiscard]] constexpr std::complex<double> times(std::complex<double> a, std::complex<double> b) noexcept {
    return {a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real()};
}

/*
One quarter turn about the y axis of a degree-l vector u_m, m = 0..l, of
complex coefficients of a real function (u_{-m} = (-1)^m conj(u_m)):
    w_m = sum_{m' = -l..l} d_{m m'} u_{m'}
        = d_{m 0} u_0 + 2 sum_{m' >= 1} d_{m m'} (Re u_{m'}   if l + m + m' is even,
                                                 i Im u_{m'} if it is odd),
using d_{m,-m'} = (-1)^{l+m} d_{m m'} of the d-matrix at pi/2, whose entries
vanish for l + m odd at m' = 0, so u_0 enters the real part only. d(m, m')
turns by +pi/2; the transposed matrix d(m', m) turns by -pi/2. The two
partial sums of each row run in increasing m', starting from d_{m 0} Re u_0
/ 2 and 0, and are doubled at the end.
*/
template <bool transposed>
constexpr void quarter_turn(std::span<const std::complex<double>> u, WignerdSpan<const double> d,
                            std::span<std::complex<double>> w) noexcept {
    const std::size_t l = u.size() - 1;
    for (std::size_t m = 0; m <= l; ++m) {
        double re = 0.5 * (transposed ? d(0, m) : d(m, 0)) * u[0].real();
        double im = 0.0;
        for (std::size_t mp = 1; mp <= l; ++mp) {
            const double dm = transposed ? d(mp, m) : d(m, mp);
            if (((l + m + mp) & 1) == 0)
                re += dm * u[mp].real();
            else
                im += dm * u[mp].imag();
        }
        w[m] = {2.0 * re, 2.0 * im};
    }
}

}  // namespace detail

/*
SO(3) rotation of one degree l = vector.size() - 1 of a complex expansion
(zest: include/wignerd_collection.hpp:113-296): a turn by gamma about z
(vector[m] *= exp_gamma[m]), a quarter turn about y, a turn by beta about z,
a quarter turn back and a turn by alpha about z. wigner_d is the d-matrix at
pi/2 of degree l; the exp_* spans hold at least l + 1 phase factors and temp
at least l + 1 elements of scratch. The rotated vector replaces vector.
*/
constexpr void rotate_l(std::span<std::complex<double>> vector, WignerdSpan<const double> wigner_d,
                        std::span<const std::complex<double>> exp_gamma,
                        std::span<const std::complex<double>> exp_beta,
                        std::span<const std::complex<double>> exp_alpha,
                        std::span<std::complex<double>> temp) noexcept {
    const std::size_t n = vector.size();
    std::span<std::complex<double>> t = temp.first(n);
    for (std::size_t m = 0; m < n; ++m) vector[m] = detail::times(vector[m], exp_gamma[m]);
    detail::quarter_turn<false>(vector, wigner_d, t);
    for (std::size_t m = 0; m < n; ++m) vector[m] = detail::times(t[m], exp_beta[m]);
    detail::quarter_turn<true>(vector, wigner_d, t);
    for (std::size_t m = 0; m < n; ++m) vector[m] = detail::times(t[m], exp_alpha[m]);
}

/*
Rotation of one degree about the z axis: vector[m] *= exp_alpha[m]
(zest: include/wignerd_collection.hpp:298-318).
*/
constexpr void polar_rotate_l(std::span<std::complex<double>> vector,
                              std::span<const std::complex<double>> exp_alpha) noexcept {
    for (std::size_t m = 0; m < vector.size(); ++m) vector[m] = detail::times(vector[m], exp_alpha[m]);
}

}  // namespace zest
