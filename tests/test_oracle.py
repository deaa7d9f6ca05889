"""Pins the oracle (oracle/zgl_oracle.c), the CPU restatement of zest's GLQ
transforms, before it is trusted as the GPU path's checker.

Every function of the oracle is checked against something it did not compute
itself: the golden fixtures written by the upstream zest library
(tests/golden/zest_golden.npz), the compiled zest library directly when it is
built (oracle/_ref), library routines (numpy leggauss, scipy lpmv and
eval_jacobi, numpy FFT), closed forms, quadrature identities, and the
known-answer transforms of zest's own tests
(tests/test_zernike_glq_transformer.cpp:57-516, tests/test_sh_glq_transformer.cpp:50-431).
"""
import math
import os

import numpy as np
import pytest
from scipy import special

import oracle
from zest_amd import synth

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "zest_golden.npz"))
CONV_ZT = [(zn, sn, sp) for zn in (0, 1) for sn in (0, 1) for sp in (-1, 1)]
CONV_ST = [(sn, sp) for sn in (0, 1) for sp in (-1, 1)]
needs_ref = pytest.mark.skipif(oracle.REF is None, reason="zest not compiled (oracle/_ref)")


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


# ---------------------------------------------------------------- GL nodes
@pytest.mark.parametrize("n", [1, 2, 3, 7, 32, 33, 69, 70, 71, 129, 136, 257, 512])
def test_gl_nodes_match_leggauss(n):
    z, w = oracle.gl_nodes_weights(n)
    zr, wr = np.polynomial.legendre.leggauss(n)
    # leggauss itself is accurate to ~1e-12 in the weights at n = 512
    assert np.abs(z - zr).max() < 5e-15
    assert rel(w, wr) < 5e-12


@pytest.mark.parametrize("n", [1, 2, 7, 32, 33, 69, 70, 71, 129, 136, 512])
def test_gl_nodes_match_zest_golden(n):
    z, w = oracle.gl_nodes_weights(n)
    # zest tabulates below 70 nodes and uses Bogaert's ~1 ulp formulas above
    assert np.abs(z - GOLDEN[f"gl_{n}_nodes"]).max() < 5e-16
    assert rel(w, GOLDEN[f"gl_{n}_weights"]) < (1e-16 if n < 70 else 1e-15)


@pytest.mark.parametrize("n", [31, 32, 111, 112])
def test_gl_invariants(n):
    """tests/test_gauss_legendre.cpp:127-210: ascending, symmetric, sum 2."""
    z, w = oracle.gl_nodes_weights(n)
    assert np.all(np.diff(z) > 0)
    assert np.array_equal(z, -z[::-1])
    assert np.array_equal(w, w[::-1])
    assert abs(w.sum() - 2.0) < 1e-14
    if n & 1:
        assert z[n // 2] == 0.0
    zp, wp = oracle.gl_packed(n)
    assert np.array_equal(zp, z[n // 2:]) and np.array_equal(wp, w[n // 2:])


# ---------------------------------------------------------------- P_lm
def plm_scipy(l, m, z, sh_norm, sh_phase):
    """Real-normalised associated Legendre function from scipy.special.lpmv
    (which carries the (-1)^m Condon-Shortley factor); zest's 'none' phase keeps
    that factor and 'cs' removes it (include/plm_recursion.hpp:197,308)."""
    norm = math.sqrt((2 - (m == 0)) * (2 * l + 1) * math.factorial(l - m) / math.factorial(l + m))
    if sh_norm == 1:
        norm /= math.sqrt(4 * math.pi)
    sign = 1.0 if sh_phase == -1 else (-1.0) ** m
    return sign * norm * special.lpmv(m, l, z)


@pytest.mark.parametrize("sh_norm,sh_phase", CONV_ST)
@pytest.mark.parametrize("z", [1.0, -1.0, 0.0, 0.97443, -0.3, 0.5])
def test_plm_matches_scipy(sh_norm, sh_phase, z):
    order = 40
    tab = oracle.plm_table(sh_norm, sh_phase, order, np.array([z]))[:, 0]
    for l in range(order):
        for m in range(l + 1):
            ref = plm_scipy(l, m, z, sh_norm, sh_phase)
            assert abs(tab[l * (l + 1) // 2 + m] - ref) < 1e-11 * max(1.0, abs(ref)), (l, m)


def test_plm_closed_forms():
    """l <= 2 closed forms (geo, phase none): P11 = -sqrt3 u, P22 = sqrt(15/4) u^2."""
    z = np.array([0.3, -0.8])
    u = np.sqrt(1 - z * z)
    t = oracle.plm_table(0, -1, 3, z)
    assert np.allclose(t[0], 1.0, rtol=0, atol=1e-15)
    assert np.allclose(t[1], math.sqrt(3) * z, rtol=1e-14)
    assert np.allclose(t[2], -math.sqrt(3) * u, rtol=1e-14)
    assert np.allclose(t[3], math.sqrt(5 / 4) * (3 * z * z - 1), rtol=1e-14)
    assert np.allclose(t[5], math.sqrt(15 / 4) * u * u, rtol=1e-14)


def test_plm_high_order_underflow_seed():
    """The 1e-280 scaling keeps high-m values finite and accurate near the poles."""
    z = np.array([0.999999, 0.0, 0.7])
    t = oracle.plm_table(0, 1, 600, z)
    assert np.all(np.isfinite(t))
    zz, ww = oracle.gl_nodes_weights(600)
    p = oracle.plm_table(0, 1, 600, zz)
    # orthogonality of P_lm over the quadrature, geo norm: int P^2 dz = 2 (2 - delta) / ...
    for (l, m) in [(599, 599), (598, 300), (450, 10)]:
        v = p[l * (l + 1) // 2 + m]
        assert abs((v * v * ww).sum() - 2.0 * (2 - (m == 0))) < 1e-10 * 2 * (2 - (m == 0))


# ---------------------------------------------------------------- R_nl
def rnl_scipy(n, l, r, normed):
    v = r ** l * special.eval_jacobi((n - l) // 2, 0.0, l + 0.5, 2 * r * r - 1)
    return v * math.sqrt(2 * n + 3) if normed else v


@pytest.mark.parametrize("zernike_norm", [0, 1])
def test_radial_matches_jacobi(zernike_norm):
    order = 48
    r = np.array([0.0, 0.3, 0.71, 1.0])
    tab = oracle.radial_table(zernike_norm, order, r)
    for n in range(order):
        for l in range(n & 1, n + 1, 2):
            ref = rnl_scipy(n, l, r, zernike_norm == 0)
            got = tab[(n + 1) ** 2 // 4 + l // 2]
            assert np.abs(got - ref).max() < 1e-11 * max(1.0, np.abs(ref).max()), (n, l)


def test_radial_endpoints_and_orthonormality():
    """R_nl(1) = 1 unnormed (tests/test_radial_zernike_recursion.cpp:456-525);
    int_0^1 R_nl R_n'l r^2 dr = delta / (2n+3) by GL quadrature (244-331)."""
    order = 30
    tab1 = oracle.radial_table(1, order, np.array([1.0]))[:, 0]
    assert np.abs(tab1 - 1.0).max() < 1e-13
    x, v = oracle.gl_nodes_weights(40)
    r = 0.5 * (1 + x)
    w = 0.5 * v * r * r
    t = oracle.radial_table(0, order, r)
    for l in range(0, order, 3):
        ns = list(range(l, order, 2))
        G = np.array([[(t[(n + 1) ** 2 // 4 + l // 2] * t[(n2 + 1) ** 2 // 4 + l // 2] * w).sum()
                       for n2 in ns] for n in ns])
        assert np.abs(G - np.eye(len(ns))).max() < 1e-13


# ---------------------------------------------------------------- DFT
@pytest.mark.parametrize("n", [1, 2, 7, 16, 127, 255])
def test_rdft_matches_numpy(n):
    x = synth.normal(n, n * 3).reshape(n, 3)
    nf = n // 2 + 1
    got = oracle.rdft_forward(x, nf, 0.5)
    ref = 0.5 * np.fft.rfft(x, axis=0)
    assert rel(got[..., 0] + 1j * got[..., 1], ref) < 1e-13
    back = oracle.rdft_backward(got / 0.5, n)
    assert rel(back, np.fft.irfft(ref / 0.5, n=n, axis=0) * n) < 1e-13


# ---------------------------------------------------------------- transforms vs zest
@pytest.mark.parametrize("order", [1, 2, 5, 6, 9])
@pytest.mark.parametrize("conv", CONV_ZT)
def test_zernike_matches_zest_golden(order, conv):
    zn, sn, sp = conv
    key = f"zt_{order}_{zn}{sn}{sp + 1}"
    if key + "_coeffs" not in GOLDEN:
        pytest.skip("convention not stored at this order")
    c, g = GOLDEN[key + "_coeffs"], GOLDEN[key + "_grid"]
    assert rel(oracle.zt_backward(c, order, zernike_norm=zn, sh_norm=sn, sh_phase=sp), GOLDEN[key + "_bwd"]) < 1e-13
    assert rel(oracle.zt_forward(g, order, zernike_norm=zn, sh_norm=sn, sh_phase=sp), GOLDEN[key + "_fwd"]) < 1e-13


def test_zernike_truncation_matches_zest_golden():
    g = GOLDEN["zt_trunc_grid"]
    assert rel(oracle.zt_forward(g, 9, 5), GOLDEN["zt_trunc_fwd5"]) < 1e-13
    out = np.full((oracle.zt_num_coeffs(12), 2), 7.25)
    assert rel(oracle.zt_forward(g, 9, 12, out=out), GOLDEN["zt_trunc_fwd12"]) < 1e-13
    assert rel(oracle.zt_backward(GOLDEN["zt_trunc_c12"], 9, 12), GOLDEN["zt_trunc_bwd12"]) < 1e-13


def test_zernike_order32_seed0_matches_zest_golden():
    """The reference's CPU-runnable case: order 32, seed 0."""
    c = GOLDEN["zt32_coeffs"]
    assert np.array_equal(c, synth.zernike_coeffs(32, 0))
    g = oracle.zt_backward(c, 32)
    idx = GOLDEN["zt32_grid_idx"]
    assert rel(g.ravel()[idx], GOLDEN["zt32_grid_val"]) < 1e-13
    sums = GOLDEN["zt32_grid_sums"]
    assert abs(g.sum() - sums[0]) < 1e-11 * abs(sums[1]) ** 0.5 * g.size ** 0.5
    assert abs((g * g).sum() / sums[1] - 1) < 1e-13
    f = oracle.zt_forward(g, 32)
    assert rel(f, GOLDEN["zt32_fwd"]) < 1e-13
    assert rel(f, c) < 1e-13  # the quadrature recovers the seeded coefficients


@pytest.mark.parametrize("seed", [0, 3])
def test_zernike_order128_matches_zest_golden(seed):
    c = synth.zernike_coeffs(128, seed)
    g = oracle.zt_backward(c, 128)
    assert rel(g.ravel()[GOLDEN[f"zt128_{seed}_grid_idx"]], GOLDEN[f"zt128_{seed}_grid_val"]) < 1e-12
    f = oracle.zt_forward(g, 128)
    assert rel(f[GOLDEN[f"zt128_{seed}_fwd_idx"]], GOLDEN[f"zt128_{seed}_fwd_val"]) < 1e-12


@pytest.mark.parametrize("order", [1, 2, 5, 7, 12])
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("conv", CONV_ST)
def test_sh_matches_zest_golden(order, layout, conv):
    sn, sp = conv
    key = f"st_{order}_{layout}_{sn}{sp + 1}"
    if key + "_coeffs" not in GOLDEN:
        pytest.skip("convention not stored at this order")
    c, g = GOLDEN[key + "_coeffs"], GOLDEN[key + "_grid"]
    assert rel(oracle.st_backward(c, order, sh_norm=sn, sh_phase=sp, layout=layout), GOLDEN[key + "_bwd"]) < 1e-13
    assert rel(oracle.st_forward(g, order, sh_norm=sn, sh_phase=sp, layout=layout), GOLDEN[key + "_fwd"]) < 1e-13


def test_sh_order512_matches_zest_golden():
    c = synth.sh_coeffs(512, 1)
    g = oracle.st_backward(c, 512)
    assert rel(g.ravel()[GOLDEN["st512_grid_idx"]], GOLDEN["st512_grid_val"]) < 1e-12
    f = oracle.st_forward(g, 512)
    assert rel(f[GOLDEN["st512_fwd_idx"]], GOLDEN["st512_fwd_val"]) < 1e-12


@needs_ref
@pytest.mark.parametrize("conv", CONV_ZT)
@pytest.mark.parametrize("order,exp_order", [(3, 3), (8, 8), (13, 13), (10, 6), (6, 10)])
def test_zernike_matches_compiled_zest(conv, order, exp_order):
    zn, sn, sp = conv
    z = oracle.ZestZernike(order, zn, sn, sp)
    g = synth.normal(order, int(np.prod(oracle.zt_shape(order)))).reshape(oracle.zt_shape(order))
    c = synth.zernike_coeffs(exp_order, order)
    out1 = np.full((oracle.zt_num_coeffs(exp_order), 2), 3.0)
    out2 = out1.copy()
    oracle.zt_forward(g, order, exp_order, zn, sn, sp, out=out1)
    z.forward(g, exp_order=exp_order, out=out2)
    assert rel(out1, out2) < 1e-13
    assert rel(oracle.zt_backward(c, order, exp_order, zn, sn, sp), z.backward(c, exp_order=exp_order)) < 1e-13


@needs_ref
@pytest.mark.parametrize("conv", CONV_ST)
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("order", [3, 8, 11, 64])
def test_sh_matches_compiled_zest(conv, layout, order):
    sn, sp = conv
    s = oracle.ZestSH(order, sn, sp, layout)
    shape = oracle.st_shape(order, layout)
    g = synth.normal(order, int(np.prod(shape))).reshape(shape)
    c = synth.sh_coeffs(order, order)
    assert rel(oracle.st_forward(g, order, sh_norm=sn, sh_phase=sp, layout=layout), s.forward(g)) < 1e-13
    assert rel(oracle.st_backward(c, order, sh_norm=sn, sh_phase=sp, layout=layout), s.backward(c)) < 1e-13


@needs_ref
@pytest.mark.parametrize("conv", CONV_ST)
@pytest.mark.parametrize("layout", [0, 1])
def test_sh_row_skipping_backward_matches_compiled_zest(conv, layout):
    """zest's row-skipping backward overload (include/sh_glq_transformer.hpp:
    680-700) on rows of a Zernike expansion equals the backward transform of
    the full expansion that oracle.st_row_skipping_expansion builds, at grid
    orders equal to, above and below the row's order, including truncations
    that change the parity of the order (zest then reads the row's data as rows
    of the other parity)."""
    sn, sp = conv
    zorder = 14
    c = synth.zernike_coeffs(zorder, 11)
    t = oracle.ZestSH(20, sn, sp, layout)
    for n, order in [(0, 1), (1, 2), (1, 20), (8, 9), (8, 16), (8, 7), (13, 14), (13, 20), (13, 12), (13, 11)]:
        e = oracle.st_row_skipping_expansion(c, zorder, n, order)
        g = oracle.st_backward(e, order, min(n + 1, order), sh_norm=sn, sh_phase=sp, layout=layout)
        assert rel(g, t.backward_zernike_row(c, zorder, n, order)) < 1e-13, (n, order)
    # the parity-changing truncation is not "drop the rows at or above the
    # grid order": that reading gives a different grid
    e = oracle.st_row_skipping_expansion(c, zorder, 13, 11)
    assert np.all(e[[l * (l + 1) // 2 for l in range(1, 11, 2)]] == 0.0)
    naive = np.zeros_like(e)
    row = c[oracle.zt_num_coeffs(13):oracle.zt_num_coeffs(14)]
    for l in range(1, 11, 2):
        naive[l * (l + 1) // 2:l * (l + 1) // 2 + l + 1] = row[l * l // 4:l * l // 4 + l + 1]
    g_naive = oracle.st_backward(naive, 11, sh_norm=sn, sh_phase=sp, layout=layout)
    assert rel(g_naive, t.backward_zernike_row(c, zorder, 13, 11)) > 1e-2


# ---------------------------------------------------------------- known answers
def ball_grid_coords(order):
    nphi, nth, nr = oracle.zt_shape(order)
    lon = 2 * np.pi * np.arange(nphi) / nphi
    z, _ = oracle.gl_nodes_weights(nth)
    x, _ = oracle.gl_nodes_weights(nr)
    return lon, z, 0.5 * (1 + x)


@pytest.mark.parametrize("conv", [(0, 0, -1), (1, 1, 1)])
@pytest.mark.parametrize("nlm", [(0, 0, 0), (2, 0, 0), (1, 1, 0), (2, 2, 1), (3, 3, -2), (5, 3, 1)])
def test_zernike_known_answer(conv, nlm):
    """A grid of one basis function R_nl P_lm cos(m lon) (sin for m < 0)
    transforms to one unit coefficient (tests/test_zernike_glq_transformer.cpp:57-516,
    run there for (normed, geo, none) and (unnormed, qm, cs))."""
    zn, sn, sp = conv
    n, l, m = nlm
    order = 6
    lon, z, r = ball_grid_coords(order)
    P = plm_scipy(l, abs(m), z, sn, sp)
    R = rnl_scipy(n, l, r, zn == 0)
    ang = np.cos(m * lon) if m >= 0 else np.sin(-m * lon)
    g = ang[:, None, None] * P[None, :, None] * R[None, None, :]
    c = oracle.zt_forward(g, order, zernike_norm=zn, sh_norm=sn, sh_phase=sp)
    expect = np.zeros_like(c)
    idx = (n + 1) * (n + 3) * (2 * n + 1) // 24 + l * l // 4 + abs(m)
    expect[idx, 0 if m >= 0 else 1] = 1.0
    assert np.abs(c - expect).max() < 1e-12
    # and the backward transform reproduces the grid (518-936)
    assert np.abs(oracle.zt_backward(expect, order, zernike_norm=zn, sh_norm=sn, sh_phase=sp) - g).max() < 1e-12


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("lm", [(0, 0), (1, 0), (2, 1), (3, 1), (4, -3)])
def test_sh_known_answer(layout, lm):
    """tests/test_sh_glq_transformer.cpp:50-431, all conventions."""
    l, m = lm
    order = 6
    for sn, sp in CONV_ST:
        shape = oracle.st_shape(order, layout)
        nlon = shape[0] if layout == 0 else shape[1]
        nlat = shape[1] if layout == 0 else shape[0]
        lon = 2 * np.pi * np.arange(nlon) / nlon
        z, _ = oracle.gl_nodes_weights(nlat)
        P = plm_scipy(l, abs(m), z, sn, sp)
        ang = np.cos(m * lon) if m >= 0 else np.sin(-m * lon)
        g = ang[:, None] * P[None, :] if layout == 0 else P[:, None] * ang[None, :]
        c = oracle.st_forward(g, order, sh_norm=sn, sh_phase=sp, layout=layout)
        expect = np.zeros_like(c)
        expect[l * (l + 1) // 2 + abs(m), 0 if m >= 0 else 1] = 1.0
        if layout == 1 and nlat & 1:
            # zest counts the equator node twice on odd latitude counts
            # (include/sh_glq_transformer.hpp:857-881); odd l+m vanish there.
            continue
        assert np.abs(c - expect).max() < 1e-12


def test_zernike_brute_force_quadrature():
    """Forward = direct triple sum of g * basis * weights (no FFT, no recursion)."""
    order = 5
    nphi, nth, nr = oracle.zt_shape(order)
    lon, z, r = ball_grid_coords(order)
    _, wz = oracle.gl_nodes_weights(nth)
    _, wr = oracle.gl_nodes_weights(nr)
    g = synth.normal(5, nphi * nth * nr).reshape(nphi, nth, nr)
    c = oracle.zt_forward(g, order, zernike_norm=1, sh_norm=1, sh_phase=1)
    W = (2 * np.pi / nphi) * wz[None, :, None] * (0.5 * wr * r * r)[None, None, :]
    for n in range(order):
        for l in range(n & 1, n + 1, 2):
            R = rnl_scipy(n, l, r, False)
            for m in range(l + 1):
                P = plm_scipy(l, m, z, 1, 1)
                base = P[None, :, None] * R[None, None, :] * W * (2 * n + 3)
                cc = (g * np.cos(m * lon)[:, None, None] * base).sum()
                ss = (g * np.sin(m * lon)[:, None, None] * base).sum()
                idx = (n + 1) * (n + 3) * (2 * n + 1) // 24 + l * l // 4 + m
                assert abs(c[idx, 0] - cc) < 1e-12 and abs(c[idx, 1] - ss) < 1e-12


# ---------------------------------------------------------------- grid evaluator
def _linspace(a, b, n):
    # tests/test_grid_evaluator.cpp:41-55 (last point exactly b)
    x = np.linspace(a, b, n)
    x[-1] = b
    return x


@pytest.mark.parametrize("conv", CONV_ST)
def test_sh_evaluate_known_answers(conv):
    """tests/test_grid_evaluator.cpp:57-281: constant, Y_10 and Y_31 + Y_4,-3 on a
    30 x 15 grid that includes both poles, tolerance 1e-13."""
    sn, sp = conv
    order = 6
    lon, col = _linspace(0, 2 * np.pi, 30), _linspace(0, np.pi, 15)
    L, C = np.meshgrid(lon, col, indexing="ij")
    z = np.cos(C)
    shnorm = 0.5 / math.sqrt(math.pi) if sn == 1 else 1.0
    phase = -1.0 if sp == -1 else 1.0
    c = np.zeros((oracle.st_num_coeffs(order), 2))
    c[0, 0] = 1.0
    assert np.abs(oracle.st_evaluate(c, order, lon, col, sn, sp) - shnorm).max() < 1e-13
    c[:] = 0
    c[1, 0] = 1.0
    assert np.abs(oracle.st_evaluate(c, order, lon, col, sn, sp) - shnorm * math.sqrt(3) * z).max() < 1e-13
    c[:] = 0
    c[3 * 4 // 2 + 1, 0] = 1.0
    c[4 * 5 // 2 + 3, 1] = 1.0
    u = np.sqrt(1 - z * z)
    f = phase * shnorm * (math.sqrt(21 / 8) * u * (5 * z * z - 1) * np.cos(L)
                          + math.sqrt(315 / 8) * u * (1 - z * z) * z * np.sin(3 * L))
    assert np.abs(oracle.st_evaluate(c, order, lon, col, sn, sp) - f).max() < 1e-13


@pytest.mark.parametrize("conv", CONV_ZT)
def test_zernike_evaluate_known_answers(conv):
    """tests/test_grid_evaluator.cpp:283-415: constant and Z_33,-2 + Z_531 on a
    14 x 7 x 7 grid including the poles, r = 0 and r = 1, tolerance 1e-13."""
    zn, sn, sp = conv
    order = 6
    lon, col, r = _linspace(0, 2 * np.pi, 14), _linspace(0, np.pi, 7), _linspace(0, 1, 7)
    L, C, R = np.meshgrid(lon, col, r, indexing="ij")
    z = np.cos(C)
    shnorm = 0.5 / math.sqrt(math.pi) if sn == 1 else 1.0
    phase = -1.0 if sp == -1 else 1.0
    c = np.zeros((oracle.zt_num_coeffs(order), 2))
    c[0, 0] = 1.0
    const = (math.sqrt(3) if zn == 0 else 1.0) * shnorm
    assert np.abs(oracle.zt_evaluate(c, order, lon, col, r, zn, sn, sp) - const).max() < 1e-13
    c[:] = 0
    tet = lambda n, l, m: (n + 1) * (n + 3) * (2 * n + 1) // 24 + l * l // 4 + m
    c[tet(3, 3, 2), 1] = 1.0
    c[tet(5, 3, 1), 0] = 1.0
    zn3 = 3.0 if zn == 0 else 1.0
    zn5 = math.sqrt(13) if zn == 0 else 1.0
    f = shnorm * (zn3 * R ** 3 * math.sqrt(105 / 4) * (1 - z * z) * z * np.sin(2 * L)
                  + zn5 * phase * (5.5 * R * R - 4.5) * R ** 3 * math.sqrt(21 / 8)
                  * np.sqrt(1 - z * z) * (5 * z * z - 1) * np.cos(L))
    assert np.abs(oracle.zt_evaluate(c, order, lon, col, r, zn, sn, sp) - f).max() < 1e-13


@pytest.mark.parametrize("order", [1, 2, 7, 16])
def test_zernike_evaluate_equals_backward_on_quadrature_nodes(order):
    """On the quadrature grid (longitudes 2 pi i / nphi, colatitudes arccos z_j,
    radial GL nodes) direct evaluation and the GLQ backward transform are the
    same sum, reached by different routes (FFT vs direct trig sums)."""
    lon, z, r = ball_grid_coords(order)
    for zn, sn, sp in [(0, 0, -1), (1, 1, 1)]:
        c = synth.zernike_coeffs(order, order + 1)
        g = oracle.zt_backward(c, order, zernike_norm=zn, sh_norm=sn, sh_phase=sp)
        v = oracle.zt_evaluate(c, order, lon, np.arccos(z), r, zn, sn, sp)
        assert rel(v, g) < 1e-13


@pytest.mark.parametrize("order", [1, 5, 16])
def test_sh_evaluate_equals_backward_on_quadrature_nodes(order):
    nlon, nlat = oracle.st_shape(order, 0)
    lon = 2 * np.pi * np.arange(nlon) / nlon
    z, _ = oracle.gl_nodes_weights(nlat)
    c = synth.sh_coeffs(order, order + 2)
    g = oracle.st_backward(c, order, sh_norm=1, sh_phase=1)
    v = oracle.st_evaluate(c, order, lon, np.arccos(z), 1, 1)
    assert rel(v, g) < 1e-13


def test_evaluate_brute_force_small():
    """Triple sum over (n, l, m) of scipy's R_nl P_lm cos/sin at random points."""
    order = 5
    rng = np.random.default_rng(7)
    lon, col, r = rng.uniform(0, 7, 4), rng.uniform(-1, 4, 3), rng.uniform(0, 1.2, 5)
    c = synth.zernike_coeffs(order, 11)
    v = oracle.zt_evaluate(c, order, lon, col, r, 1, 0, -1)
    ref = np.zeros_like(v)
    z = np.cos(col)
    for n in range(order):
        for l in range(n & 1, n + 1, 2):
            R = rnl_scipy(n, l, r, False)
            for m in range(l + 1):
                # scipy's P_l^m carries |sin|, the recursion sqrt((1-z)(1+z)): equal
                P = plm_scipy(l, m, z, 0, -1)
                idx = (n + 1) * (n + 3) * (2 * n + 1) // 24 + l * l // 4 + m
                ang = c[idx, 0] * np.cos(m * lon) + c[idx, 1] * np.sin(m * lon)
                ref += ang[:, None, None] * P[None, :, None] * R[None, None, :]
    assert rel(v, ref) < 1e-12


def test_evaluate_empty_and_order0():
    c = np.zeros((1, 2))
    assert oracle.zt_evaluate(c, 0, [0.1], [0.2], [0.3]).tolist() == [[[0.0]]]
    assert oracle.zt_evaluate(np.zeros((oracle.zt_num_coeffs(3), 2)), 3, [], [0.2], [0.3]).shape == (0, 1, 1)
    assert oracle.st_evaluate(c, 0, [0.1, 0.2], [0.2]).tolist() == [[0.0], [0.0]]


@needs_ref
@pytest.mark.parametrize("conv", CONV_ZT)
@pytest.mark.parametrize("order", [1, 3, 10, 33])
def test_zernike_evaluate_matches_compiled_zest(conv, order):
    zn, sn, sp = conv
    rng = np.random.default_rng(order)
    lon = np.concatenate([[0.0, 2 * np.pi], rng.uniform(-1, 8, 9)])
    col = np.concatenate([[0.0, np.pi], rng.uniform(0, np.pi, 6)])
    r = np.concatenate([[0.0, 1.0], rng.uniform(0, 1.1, 4)])
    c = synth.zernike_coeffs(order, 5 + order)
    assert rel(oracle.zt_evaluate(c, order, lon, col, r, zn, sn, sp),
               oracle.ref_zt_evaluate(c, order, lon, col, r, zn, sn, sp)) < 1e-13


@needs_ref
@pytest.mark.parametrize("conv", CONV_ST)
@pytest.mark.parametrize("order", [1, 2, 17, 100])
def test_sh_evaluate_matches_compiled_zest(conv, order):
    sn, sp = conv
    rng = np.random.default_rng(order + 1)
    lon = np.concatenate([[0.0, 2 * np.pi], rng.uniform(-1, 8, 13)])
    col = np.concatenate([[0.0, np.pi], rng.uniform(0, np.pi, 10)])
    c = synth.sh_coeffs(order, 2 + order)
    assert rel(oracle.st_evaluate(c, order, lon, col, sn, sp),
               oracle.ref_st_evaluate(c, order, lon, col, sn, sp)) < 1e-13


# ---------------------------------------------------------------- power spectra
@needs_ref
@pytest.mark.parametrize("zn", [0, 1])
@pytest.mark.parametrize("sn", [0, 1])
@pytest.mark.parametrize("order", [1, 2, 9, 40])
def test_zernike_power_spectrum_matches_compiled_zest(zn, sn, order):
    """zest writes the (n, l < n) entries (its loop bound skips l = n)."""
    c = synth.zernike_coeffs(order, order)
    ref = oracle.ref_zt_power_spectrum(c, order, zn, sn)
    got = oracle.zt_power_spectrum(c, order, sn)
    written = ~np.isnan(ref)
    assert written.sum() == (order + 1) ** 2 // 4 - order
    if written.any():
        assert np.abs(got[written] - ref[written]).max() <= 1e-15 * np.abs(ref[written]).max()


@pytest.mark.parametrize("sn", [0, 1])
def test_power_spectrum_known_answers(sn):
    norm = 1.0 if sn == 1 else 1 / (4 * math.pi)
    c = np.zeros((oracle.st_num_coeffs(5), 2))
    c[3 * 4 // 2 + 2] = (0.6, 0.8)
    c[2 * 3 // 2 + 0] = (0.5, 7.0)  # the sin part at m = 0 does not count
    p = oracle.st_power_spectrum(c, 5, sn)
    assert np.allclose(p, norm * np.array([0, 0, 0.25, 1.0, 0]), rtol=0, atol=1e-16)
    z = np.zeros((oracle.zt_num_coeffs(5), 2))
    tet = lambda n, l, m: (n + 1) * (n + 3) * (2 * n + 1) // 24 + l * l // 4 + m
    z[tet(4, 4, 3)] = (3.0, 4.0)
    z[tet(4, 4, 0)] = (1.0, 9.0)
    z[tet(3, 1, 1)] = (2.0, 0.0)
    p = oracle.zt_power_spectrum(z, 5, sn)
    expect = np.zeros_like(p)
    expect[(5 * 5) // 4 + 2] = 26.0
    expect[(4 * 4) // 4 + 0] = 4.0
    assert np.allclose(p, norm * expect, rtol=0, atol=1e-15)
    # cross spectrum: bilinear, symmetric, and equal to the power spectrum on the diagonal
    a, b = synth.sh_coeffs(7, 1), synth.sh_coeffs(7, 2)
    pab = oracle.st_power_spectrum(a, 7, sn, b)
    assert np.allclose(pab, oracle.st_power_spectrum(b, 7, sn, a), rtol=1e-15, atol=0)
    assert np.allclose(oracle.st_power_spectrum(a + b, 7, sn),
                       oracle.st_power_spectrum(a, 7, sn) + 2 * pab + oracle.st_power_spectrum(b, 7, sn),
                       rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("sn", [0, 1])
def test_sh_power_spectrum_parseval(sn):
    """sum_l P(l) = int f^2 dOmega / (4 pi for geo: one more 4 pi from the
    basis norm), by quadrature of the oracle's backward grid."""
    L = 12
    c = synth.sh_coeffs(L, 5)
    c[[l * (l + 1) // 2 for l in range(L)], 1] = 0.0
    g = oracle.st_backward(c, L, sh_norm=sn, sh_phase=1)
    nlon, nlat = oracle.st_shape(L, 0)
    _, w = oracle.gl_nodes_weights(nlat)
    integral = float((g * g * w[None, :]).sum() * 2 * math.pi / nlon)
    total = oracle.st_power_spectrum(c, L, sn).sum()
    expect = integral if sn == 1 else integral / (4 * math.pi) ** 2
    assert abs(total - expect) <= 1e-13 * expect


@pytest.mark.parametrize("sn", [0, 1])
def test_zernike_power_spectrum_parseval(sn):
    """Normed Zernike functions are orthonormal in the ball (qm) or carry 4 pi
    (geo): sum_nl P(n, l) = int f^2 dV (/ (4 pi)^2 for geo)."""
    N = 10
    c = synth.zernike_coeffs(N, 6)
    g = oracle.zt_backward(c, N, zernike_norm=0, sh_norm=sn, sh_phase=-1)
    nphi, nth, nr = oracle.zt_shape(N)
    _, wz = oracle.gl_nodes_weights(nth)
    x, wx = oracle.gl_nodes_weights(nr)
    r = 0.5 * (1 + x)
    W = (2 * math.pi / nphi) * wz[None, :, None] * (0.5 * wx * r * r)[None, None, :]
    integral = float((g * g * W).sum())
    total = oracle.zt_power_spectrum(c, N, sn).sum()
    expect = integral if sn == 1 else integral / (4 * math.pi) ** 2
    assert abs(total - expect) <= 1e-13 * expect


# ---------------------------------------------------------------- rotations
def _wigner_textbook(j, mp, m, beta):
    """Wigner's explicit sum for d^j_{m',m}(beta) at beta = pi/2, where every
    cos^a sin^b (a + b = 2j) is 2^-j: the alternating sum is exact in rationals."""
    from fractions import Fraction
    assert beta == math.pi / 2
    F = math.factorial
    s = Fraction(0)
    for k in range(0, 2 * j + 1):
        a = (j + m - k, k, mp - m + k, j - mp - k)
        if min(a) < 0:
            continue
        s += Fraction((-1) ** (mp - m + k), F(a[0]) * F(a[1]) * F(a[2]) * F(a[3]))
    return float(s) * math.sqrt(F(j + mp) * F(j - mp) * F(j + m) * F(j - m)) / 2.0 ** j


def test_wignerd_table_textbook_and_orthogonal():
    order = 12
    d = oracle.wignerd_table(order)
    for l in range(order):
        base = l * (l + 1) * (2 * l + 1) // 6
        for m1 in range(l + 1):
            for m2 in range(l + 1):
                assert abs(d[base + (l + 1) * m1 + m2] - _wigner_textbook(l, m1, m2, math.pi / 2)) < 1e-14
    # high order: the full (2l+1)^2 matrix from the stored quarter is orthogonal
    order = 200
    d = oracle.wignerd_table(order)
    for l in (1, 2, 57, 199):
        q = d[l * (l + 1) * (2 * l + 1) // 6:][:(l + 1) ** 2].reshape(l + 1, l + 1)
        D = np.zeros((2 * l + 1, 2 * l + 1))
        # d_{a,b} = (-1)^(b-a) d_{b,a} = d_{-b,-a};  d_{a,-b}(pi/2) = (-1)^(l+a) d_{a,b}(pi/2)
        for a in range(-l, l + 1):
            for b in range(-l, l + 1):
                if a >= 0:
                    D[a + l, b + l] = q[a, b] if b >= 0 else (-1) ** (l + a) * q[a, -b]
                elif b <= 0:
                    D[a + l, b + l] = q[-b, -a]
                else:
                    D[a + l, b + l] = (-1) ** (b - a) * (-1) ** (l + b) * q[b, -a]
        if l <= 2:
            for a in range(-l, l + 1):
                for b in range(-l, l + 1):
                    assert abs(D[a + l, b + l] - _wigner_textbook(l, a, b, math.pi / 2)) < 1e-14
        assert np.abs(D @ D.T - np.eye(2 * l + 1)).max() < 1e-12, l


@needs_ref
@pytest.mark.parametrize("order", [1, 2, 33, 128])
def test_wignerd_table_matches_compiled_zest(order):
    a, b = oracle.wignerd_table(order), oracle.ref_wignerd_table(order)
    assert np.abs(a - b).max() < 1e-14


def _rot(a, b, g):
    Rz = lambda t: np.array([[math.cos(t), -math.sin(t), 0], [math.sin(t), math.cos(t), 0], [0, 0, 1]])
    Ry = lambda t: np.array([[math.cos(t), 0, math.sin(t)], [0, 1, 0], [-math.sin(t), 0, math.cos(t)]])
    return Rz(a) @ Ry(b) @ Rz(g)


@pytest.mark.parametrize("rtype", [0, 1])
def test_rotation_is_rotation_of_the_function(rtype):
    """Phase convention none: the rotated expansion evaluated at x equals the
    original at R^T x (object) or R x (coordinate), R = Rz(alpha) Ry(beta)
    Rz(gamma), intrinsic ZYZ (include/rotor.hpp:104-122)."""
    N = 7
    ang = (0.3, 1.1, -2.0)
    R = _rot(*ang)
    M = R.T if rtype == 0 else R
    rng = np.random.default_rng(4)
    pts = [(rng.uniform(0, 2 * np.pi), rng.uniform(0, np.pi), rng.uniform(0.1, 1)) for _ in range(12)]
    c = synth.sh_coeffs(N, 1)
    z = synth.zernike_coeffs(N, 2)
    cr = oracle.rotate(c, N, oracle.SH_KIND, ang, rtype)
    zr = oracle.rotate(z, N, oracle.ZERNIKE_KIND, ang, rtype)
    for lon, col, r in pts:
        x = np.array([math.sin(col) * math.cos(lon), math.sin(col) * math.sin(lon), math.cos(col)])
        y = M @ x
        col2, lon2 = math.acos(max(-1.0, min(1.0, y[2]))), math.atan2(y[1], y[0])
        a = oracle.st_evaluate(cr, N, [lon], [col], 0, -1)[0, 0]
        b = oracle.st_evaluate(c, N, [lon2], [col2], 0, -1)[0, 0]
        assert abs(a - b) < 1e-13
        a = oracle.zt_evaluate(zr, N, [lon], [col], [r], 1, 0, -1)[0, 0, 0]
        b = oracle.zt_evaluate(z, N, [lon2], [col2], [r], 1, 0, -1)[0, 0, 0]
        assert abs(a - b) < 1e-13


def test_rotation_invariants():
    N = 20
    c = synth.sh_coeffs(N, 8)
    z = synth.zernike_coeffs(N, 9)
    ang = (2.5, 0.4, 1.3)
    # degree-wise power is invariant
    for kind, x, pw in ((oracle.SH_KIND, c, oracle.st_power_spectrum), (oracle.ZERNIKE_KIND, z, oracle.zt_power_spectrum)):
        r = oracle.rotate(x, N, kind, ang, 0)
        assert np.allclose(pw(r, N, 1), pw(x, N, 1), rtol=1e-12, atol=1e-14)
        # object rotation followed by the coordinate rotation of the same angles is the identity
        back = oracle.rotate(r, N, kind, ang, 1)
        assert np.abs(back - x).max() < 1e-12
        # beta = 0: a z rotation by alpha + gamma, the same as polar_rotate
        p = oracle.rotate(x, N, kind, (0.7, 0.0, 0.4), 0)
        q = oracle.polar_rotate(x, N, kind, 1.1, 0)
        assert np.abs(p - q).max() < 1e-12


@needs_ref
@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("order", [1, 2, 9, 64])
@pytest.mark.parametrize("rtype", [0, 1])
def test_rotation_matches_compiled_zest(kind, order, rtype):
    c = synth.zernike_coeffs(order, order) if kind == 0 else synth.sh_coeffs(order, order)
    for ang in [(0.3, 1.1, -2.0), (4.0, 2.9, 0.0)]:
        a = oracle.rotate(c, order, kind, ang, rtype)
        b = oracle.ref_rotate(c, order, kind, ang, rtype)
        assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()
    a = oracle.polar_rotate(c, order, kind, 0.9, rtype)
    b = oracle.ref_polar_rotate(c, order, kind, 0.9, rtype)
    assert np.abs(a - b).max() <= 1e-15 * np.abs(b).max()
