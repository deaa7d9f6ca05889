// zest::rotate_l and zest::polar_rotate_l through the drop-in header.
//
// zest's per-degree rotation building blocks (include/wignerd_collection.hpp:
// 122, 304) are public functions that callers use to rotate single degrees of
// a complex expansion themselves. This program
//  1. applies the drop-in's rotate_l / polar_rotate_l
//     (include/zest_amd/wignerd_collection.hpp, compiled as zest's
//     wignerd_collection.hpp) to random degree-l vectors with random phase
//     factors and writes each case to <out_dir>/rl_<k>.bin, for comparison
//     with zest's own functions on the CPU (tests/test_dropin_gpu.py):
//       int64 header[2] = {l, 0}
//       complex vector[l+1], exp_gamma[l+1], exp_beta[l+1], exp_alpha[l+1],
//       rotate_l result[l+1], polar_rotate_l(vector, exp_alpha) result[l+1]
//     (complex values as interleaved re, im doubles);
//  2. rotates real SH and Zernike expansions the way zest's Rotor does
//     (include/rotor.hpp:124-240, 310-335: convert to the complex form, phase
//     factors exp(-i m angle) of the converted Euler angles, rotate_l or
//     polar_rotate_l on every degree l >= 1, convert back) with the drop-in's
//     functions and a WignerdPiHalfCollection from the GPU, and checks the
//     result against the drop-in Rotor, which runs the whole rotation on the
//     GPU, at 1e-12 relative to the largest coefficient;
//  3. checks that zero angles give the identity (the quarter turn and its
//     transpose cancel) at 1e-13.
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <numbers>
#include <span>
#include <string>
#include <vector>

#include "real_sh_expansion.hpp"
#include "rotor.hpp"
#include "wignerd_collection.hpp"
#include "zernike_expansion.hpp"

namespace {

using C = std::complex<double>;
using zest::st::SHNorm;
using zest::st::SHPhase;

std::uint64_t g_state = 0x2545f4914f6cdd1dull;

double next_uniform() {
    // splitmix64, mapped to [-1, 1)
    std::uint64_t z = (g_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return double(z >> 11) * 0x1.0p-52 - 1.0;
}

int g_failures = 0;

void check(bool ok, const char* what, double err) {
    if (!ok) {
        std::printf("FAILED %s (error %.3e)\n", what, err);
        ++g_failures;
    }
}

std::vector<C> phases(double angle, std::size_t n) {
    std::vector<C> e(n);
    for (std::size_t m = 0; m < n; ++m) e[m] = std::polar(1.0, -double(m) * angle);
    return e;
}

bool write_degree_cases(const std::string& dir, const zest::WignerdPiHalfCollection& wig) {
    const std::size_t degrees[] = {0, 1, 2, 3, 4, 7, 16, 33, 64, 127};
    int k = 0;
    for (std::size_t l : degrees) {
        const std::size_t n = l + 1;
        std::vector<C> v(n), eg, eb, ea, temp(n + 3);
        for (std::size_t m = 0; m < n; ++m) v[m] = {next_uniform(), m ? next_uniform() : 0.0};
        eg = phases(3.0 * next_uniform(), n);
        eb = phases(3.0 * next_uniform(), n);
        ea = phases(3.0 * next_uniform(), n);
        std::vector<C> r = v, p = v;
        zest::rotate_l(r, wig[l], eg, eb, ea, temp);
        zest::polar_rotate_l(p, ea);
        const std::string path = dir + "/rl_" + std::to_string(k++) + ".bin";
        std::FILE* f = std::fopen(path.c_str(), "wb");
        if (!f) return false;
        const std::int64_t head[2] = {std::int64_t(l), 0};
        bool ok = std::fwrite(head, sizeof head, 1, f) == 1;
        for (const auto* a : {&v, &eg, &eb, &ea, &r, &p})
            ok = ok && std::fwrite(a->data(), sizeof(C), n, f) == n;
        if (std::fclose(f) != 0 || !ok) return false;
    }
    return true;
}

// zest's Euler angle conversion (include/rotor.hpp:68-81)
std::array<double, 3> rotation_angles(const std::array<double, 3>& a, zest::RotationType type) {
    const double ar = a[0] - 0.5 * std::numbers::pi, br = a[1], gr = a[2] + 0.5 * std::numbers::pi;
    if (type == zest::RotationType::object) return {ar, br, gr};
    return {-gr, -br, -ar};
}

double rel_diff(std::span<const std::array<double, 2>> a, std::span<const std::array<double, 2>> b) {
    double num = 0.0, den = 0.0;
    for (std::size_t i = 0; i < a.size(); ++i)
        for (int j = 0; j < 2; ++j) {
            num = std::max(num, std::fabs(a[i][j] - b[i][j]));
            den = std::max(den, std::fabs(b[i][j]));
        }
    return num / den;
}

template <SHNorm sn, SHPhase sp>
void check_sh_rotor(const zest::WignerdPiHalfCollection& wig, zest::RotationType type) {
    constexpr std::size_t order = 48;
    zest::st::RealSHExpansion<sn, sp> a(order);
    for (auto& c : a.flatten()) c = {next_uniform(), next_uniform()};
    for (auto l : a.indices()) a(l, 0)[1] = 0.0;
    zest::st::RealSHExpansion<sn, sp> b = a;
    const std::array<double, 3> angles = {0.7, 2.1, -1.3};

    zest::Rotor rotor(order);
    rotor.rotate(b, wig, angles, type);

    const auto [ar, br, gr] = rotation_angles(angles, type);
    const auto ea = phases(ar, order), eb = phases(br, order), eg = phases(gr, order);
    std::vector<C> temp(order);
    auto cx = zest::st::to_complex_expansion<sn, sp>(a);
    for (auto l : a.indices(1)) zest::rotate_l(cx[l], wig[l], eg, eb, ea, temp);
    zest::st::to_real_expansion<sn, sp>(cx);
    const double e = rel_diff(a.flatten(), b.flatten());
    check(e < 1e-12, "SH rotation composed of rotate_l vs Rotor::rotate", e);
    std::printf("SH order %zu: rotate_l vs Rotor %.2e\n", order, e);
}

template <SHNorm sn, SHPhase sp>
void check_sh_polar(zest::RotationType type) {
    constexpr std::size_t order = 40;
    zest::st::RealSHExpansion<sn, sp> a(order);
    for (auto& c : a.flatten()) c = {next_uniform(), next_uniform()};
    for (auto l : a.indices()) a(l, 0)[1] = 0.0;
    zest::st::RealSHExpansion<sn, sp> b = a;
    const double angle = -0.9;
    zest::Rotor rotor(order);
    rotor.polar_rotate(b, angle, type);
    const double s = type == zest::RotationType::object ? 1.0 : -1.0;
    const auto ea = phases(s * angle, order);
    auto cx = zest::st::to_complex_expansion<sn, sp>(a);
    for (auto l : a.indices(1)) zest::polar_rotate_l(cx[l], ea);
    zest::st::to_real_expansion<sn, sp>(cx);
    const double e = rel_diff(a.flatten(), b.flatten());
    check(e < 1e-12, "SH rotation composed of polar_rotate_l vs Rotor::polar_rotate", e);
}

void check_zernike_rotor(const zest::WignerdPiHalfCollection& wig, zest::RotationType type) {
    using zest::zt::ZernikeNorm;
    constexpr std::size_t order = 24;
    constexpr ZernikeNorm zn = ZernikeNorm::normed;
    constexpr SHNorm sn = SHNorm::geo;
    constexpr SHPhase sp = SHPhase::none;
    zest::zt::RealZernikeExpansion<zn, sn, sp> a(order);
    for (auto& c : a.flatten()) c = {next_uniform(), next_uniform()};
    for (auto n : a.indices())
        for (auto l : a[n].indices()) a[n][l][0][1] = 0.0;
    zest::zt::RealZernikeExpansion<zn, sn, sp> b = a;
    const std::array<double, 3> angles = {-2.4, 0.45, 1.9};
    zest::Rotor rotor(order);
    rotor.rotate(b, wig, angles, type);

    const auto [ar, br, gr] = rotation_angles(angles, type);
    const auto ea = phases(ar, order), eb = phases(br, order), eg = phases(gr, order);
    std::vector<C> temp(order);
    auto cx = zest::zt::to_complex_expansion<zn, sn, sp>(a);
    for (auto n : a.indices(1)) {
        auto cn = cx[n];
        for (auto l : cn.indices(1)) zest::rotate_l(cn[l], wig[l], eg, eb, ea, temp);
    }
    zest::zt::to_real_expansion<zn, sn, sp>(cx);
    const double e = rel_diff(a.flatten(), b.flatten());
    check(e < 1e-12, "Zernike rotation composed of rotate_l vs Rotor::rotate", e);
    std::printf("Zernike order %zu: rotate_l vs Rotor %.2e\n", order, e);
}

void check_identity(const zest::WignerdPiHalfCollection& wig) {
    for (std::size_t l : {1, 2, 5, 6, 63, 64}) {
        const std::size_t n = l + 1;
        std::vector<C> v(n), temp(n);
        for (std::size_t m = 0; m < n; ++m) v[m] = {next_uniform(), m ? next_uniform() : 0.0};
        const std::vector<C> ones(n, C{1.0, 0.0});
        std::vector<C> r = v;
        zest::rotate_l(r, wig[l], ones, ones, ones, temp);
        double err = 0.0;
        for (std::size_t m = 0; m < n; ++m) err = std::max(err, std::abs(r[m] - v[m]));
        check(err < 1e-13, "rotate_l with zero angles is the identity", err);
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <out_dir>\n", argv[0]);
        return 2;
    }
    const zest::WignerdPiHalfCollection wig(128);
    if (!write_degree_cases(argv[1], wig)) {
        std::fprintf(stderr, "cannot write to %s\n", argv[1]);
        return 2;
    }
    check_identity(wig);
    for (auto type : {zest::RotationType::object, zest::RotationType::coordinate}) {
        check_sh_rotor<SHNorm::geo, SHPhase::none>(wig, type);
        check_sh_rotor<SHNorm::qm, SHPhase::cs>(wig, type);
        check_sh_polar<SHNorm::geo, SHPhase::none>(type);
        check_sh_polar<SHNorm::qm, SHPhase::cs>(type);
        check_zernike_rotor(wig, type);
    }
    std::printf("%d failures\n", g_failures);
    return g_failures ? 1 : 0;
}
