// Drop-in replacement of zest's "wignerd_collection.hpp" (sebsassi/zest,
// include/wignerd_collection.hpp) backed by the MI355X library
// (include/zgl.h): WignerdSpan and WignerdPiHalfCollection keep zest's
// names, accessors and table layout; the table is generated on the GPU
// (zgl_wignerd_pi2) and held on the host, as zest holds it. zest's
// per-degree building blocks rotate_l / polar_rotate_l are not provided: the
// drop-in Rotor (rotor.hpp) runs whole rotations on the GPU.
#pragma once

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

}  // namespace zest
