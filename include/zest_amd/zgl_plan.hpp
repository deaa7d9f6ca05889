// RAII owner of a zgl_plan (include/zgl.h) for the C++20 drop-in headers.
// Errors become exceptions, as in zest: allocation failures throw
// std::bad_alloc (zest's AlignedAllocator, include/alignment.hpp:469-479),
// everything else std::runtime_error with the library's message.
#pragma once

#include <cstddef>
#include <new>
#include <stdexcept>
#include <utility>

#include "../zgl.h"

namespace zest_amd {

inline void zgl_check(zgl_status s) {
    if (s == ZGL_SUCCESS) return;
    if (s == ZGL_ERROR_OUT_OF_MEMORY) throw std::bad_alloc();
    throw std::runtime_error(zgl_last_error_string());
}

class Plan {
public:
    Plan(zgl_kind kind, int zernike_norm, int sh_norm, int sh_phase, int layout,
         std::size_t order, zgl_precision precision = ZGL_FP64)
        : m_kind(kind), m_zn(zernike_norm), m_sn(sh_norm), m_sp(sh_phase), m_layout(layout),
          m_precision(precision) {
        zgl_check(zgl_plan_create(&m_plan, kind, zernike_norm, sh_norm, sh_phase, layout, order,
                                  precision));
    }
    ~Plan() { zgl_plan_destroy(m_plan); }

    // Copies get their own plan (zest's transformers copy their tables).
    Plan(const Plan& o) : Plan(o.m_kind, o.m_zn, o.m_sn, o.m_sp, o.m_layout, o.order(), o.m_precision) {}
    Plan& operator=(const Plan& o) {
        if (this != &o) {
            Plan tmp(o);
            std::swap(*this, tmp);
        }
        return *this;
    }
    Plan(Plan&& o) noexcept
        : m_plan(std::exchange(o.m_plan, nullptr)), m_kind(o.m_kind), m_zn(o.m_zn), m_sn(o.m_sn),
          m_sp(o.m_sp), m_layout(o.m_layout), m_precision(o.m_precision) {}
    Plan& operator=(Plan&& o) noexcept {
        std::swap(m_plan, o.m_plan);
        m_kind = o.m_kind;
        m_zn = o.m_zn;
        m_sn = o.m_sn;
        m_sp = o.m_sp;
        m_layout = o.m_layout;
        m_precision = o.m_precision;
        return *this;
    }

    [[nodiscard]] std::size_t order() const noexcept { return zgl_plan_order(m_plan); }
    void resize(std::size_t order) { zgl_check(zgl_plan_resize(m_plan, order)); }

    // Synchronous host-memory transforms of one instance.
    void forward(const double* grid, std::size_t grid_order, double* coeffs, std::size_t exp_order) {
        zgl_check(zgl_forward(m_plan, grid, grid_order, coeffs, exp_order, 1, 0, 0, ZGL_MEM_HOST,
                              nullptr));
    }
    void backward(const double* coeffs, std::size_t exp_order, double* grid, std::size_t grid_order) {
        zgl_check(zgl_backward(m_plan, coeffs, exp_order, grid, grid_order, 1, 0, 0, ZGL_MEM_HOST,
                               nullptr));
    }

    [[nodiscard]] zgl_plan* get() const noexcept { return m_plan; }

private:
    zgl_plan* m_plan = nullptr;
    zgl_kind m_kind;
    int m_zn, m_sn, m_sp, m_layout;
    zgl_precision m_precision;
};

}  // namespace zest_amd
