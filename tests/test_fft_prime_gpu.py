"""Longitude FFT lengths with a prime factor above 17 (2N-1 = 19, 31, 59 * 3,
37 * 5, 127, 191, 199): those passes run on the matrix cores as two real
products over the symmetric sums and differences of the inputs. Checked
against the FP64 oracle at 1e-12, against the VALU pass (ZGL_FFT_MFMA=0)
to rounding level, in the FP32 mode against the oracle at the FP32 level,
and for independence from the tile width (bit-identical; widths 16 and 8 run
the lean kernel variant, widths 32 and 4 the general one)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _roundtrip(device, kind, order, c, layout=0, precision=0):
    import torch
    import zest_amd
    zt = kind == "zernike"
    plan = zest_amd.Plan(zest_amd.ZERNIKE if zt else zest_amd.SH, order, layout=layout,
                         precision=precision)
    shape = zest_amd.zt_shape(order) if zt else zest_amd.st_shape(order, layout)
    batch = c.shape[0]
    dc = torch.from_numpy(c).to(device)
    dg = torch.empty((batch,) + shape, dtype=torch.float64, device=device)
    co = torch.zeros_like(dc)
    plan.backward(dc, dg, order, batch=batch)
    plan.forward(dg, co, order, batch=batch)
    torch.cuda.synchronize()
    return dg.cpu().numpy(), co.cpu().numpy()


# 2N-1: 19, 31, 127, 3 * 59, 5 * 37, 191 and 11 * 19 (the general kernel
# variant; the others run the lean one, whose other radices are at most 7);
# SH 2L-1 = 31, 199 and 437 = 23 * 19 (two matrix-core passes, the second
# with input twiddles); SH 263 and 383, primes above 255 (the 512-thread
# variant)
CASES = [("zernike", 10), ("zernike", 16), ("zernike", 64), ("zernike", 89), ("zernike", 93),
         ("zernike", 96), ("zernike", 105), ("sh", 16), ("sh", 100), ("sh", 219),
         ("sh", 132), ("sh", 192)]


@pytest.mark.parametrize("kind,order", CASES)
def test_prime_radix_fp64(device, monkeypatch, kind, order):
    import oracle
    from zest_amd import synth
    zt = kind == "zernike"
    gen = synth.zernike_coeffs if zt else synth.sh_coeffs
    c = np.stack([gen(order, 300 + b) for b in range(2)])
    g_mfma, c_mfma = _roundtrip(device, kind, order, c)
    monkeypatch.setenv("ZGL_FFT_MFMA", "0")
    g_valu, c_valu = _roundtrip(device, kind, order, c)
    for b in range(2):
        g_ref = oracle.zt_backward(c[b], order) if zt else oracle.st_backward(c[b], order)
        c_ref = oracle.zt_forward(g_ref, order) if zt else oracle.st_forward(g_ref, order)
        eg, ec = _rel(g_mfma[b], g_ref), _rel(c_mfma[b], c_ref)
        print(f"{kind} {order}: backward {eg:.2e}, forward {ec:.2e}, "
              f"vs VALU pass {_rel(g_mfma[b], g_valu[b]):.2e}")
        assert eg <= 1e-12 and ec <= 1e-12
        assert _rel(g_mfma[b], g_valu[b]) <= 1e-13 and _rel(c_mfma[b], c_valu[b]) <= 1e-13


@pytest.mark.parametrize("kind,order", [("zernike", 64), ("zernike", 93), ("sh", 100)])
def test_prime_radix_fp32(device, kind, order):
    import oracle
    import zest_amd
    from zest_amd import synth
    zt = kind == "zernike"
    gen = synth.zernike_coeffs if zt else synth.sh_coeffs
    c = np.stack([gen(order, 400 + b) for b in range(2)])
    g, co = _roundtrip(device, kind, order, c, precision=zest_amd.FP32)
    for b in range(2):
        g_ref = oracle.zt_backward(c[b], order) if zt else oracle.st_backward(c[b], order)
        c_ref = oracle.zt_forward(g_ref, order) if zt else oracle.st_forward(g_ref, order)
        print(f"{kind} {order} fp32: backward {_rel(g[b], g_ref):.2e}, forward {_rel(co[b], c_ref):.2e}")
        assert _rel(g[b], g_ref) < 1e-6 and _rel(co[b], c_ref) < 1e-6


@pytest.mark.parametrize("kind,order,layout", [("zernike", 96, 0), ("sh", 100, 1)])
def test_prime_radix_tile_widths(device, monkeypatch, kind, order, layout):
    from zest_amd import synth
    zt = kind == "zernike"
    gen = synth.zernike_coeffs if zt else synth.sh_coeffs
    c = np.stack([gen(order, 500 + b) for b in range(2)])
    res = {}
    for tile in ("32", "16", "8", "4"):
        monkeypatch.setenv("ZGL_FFT_TILE", tile)
        res[tile] = _roundtrip(device, kind, order, c, layout=layout)
    for tile in ("32", "8", "4"):
        assert np.array_equal(res[tile][0], res["16"][0]), tile
        assert np.array_equal(res[tile][1], res["16"][1]), tile


@pytest.mark.parametrize("order,layout", [(512, 0), (512, 1), (300, 0)])
def test_long_lengths_global_twiddles(device, monkeypatch, order, layout):
    """FFT lengths from 384 on keep their twiddle tables in global memory
    (2L-1 = 1023 = 3 * 11 * 31; 599, a prime above the matrix-core range, on
    the VALU pass): SH against the oracle at 1e-12, and tile widths 8 and 4
    bit-identical."""
    import oracle
    from zest_amd import synth
    c = np.stack([synth.sh_coeffs(order, 600 + b) for b in range(2)])
    res = {}
    for tile in ("8", "4"):
        monkeypatch.setenv("ZGL_FFT_TILE", tile)
        res[tile] = _roundtrip(device, "sh", order, c, layout=layout)
    assert np.array_equal(res["8"][0], res["4"][0]) and np.array_equal(res["8"][1], res["4"][1])
    g, co = res["4"]
    for b in range(2):
        g_ref = oracle.st_backward(c[b], order, layout=layout)
        c_ref = oracle.st_forward(g_ref, order, layout=layout)
        print(f"sh {order} layout {layout}: backward {_rel(g[b], g_ref):.2e}, forward {_rel(co[b], c_ref):.2e}")
        assert _rel(g[b], g_ref) <= 1e-12 and _rel(co[b], c_ref) <= 1e-12


@pytest.mark.parametrize("kind,order,layout", [("zernike", 20, 0), ("zernike", 33, 0),
                                               ("sh", 41, 0), ("sh", 41, 1)])
def test_fp32_tile_widths(device, monkeypatch, kind, order, layout):
    """FP32 plans without a matrix-core pass (2N-1 = 39 = 3 * 13, 65 = 5 * 13,
    81 = 3^4): width 32 runs 512 threads per workgroup, the others 256. All
    widths give bit-identical results, and width 32 matches the oracle at the
    FP32 level."""
    import oracle
    import zest_amd
    from zest_amd import synth
    zt = kind == "zernike"
    gen = synth.zernike_coeffs if zt else synth.sh_coeffs
    c = np.stack([gen(order, 700 + b) for b in range(3)])
    res = {}
    for tile in ("32", "16", "8"):
        monkeypatch.setenv("ZGL_FFT_TILE", tile)
        res[tile] = _roundtrip(device, kind, order, c, layout=layout, precision=zest_amd.FP32)
    for tile in ("32", "8"):
        assert np.array_equal(res[tile][0], res["16"][0]), tile
        assert np.array_equal(res[tile][1], res["16"][1]), tile
    g, co = res["32"]
    for b in range(3):
        if zt:
            g_ref = oracle.zt_backward(c[b], order)
            c_ref = oracle.zt_forward(g_ref, order)
        else:
            g_ref = oracle.st_backward(c[b], order, layout=layout)
            c_ref = oracle.st_forward(g_ref, order, layout=layout)
        print(f"{kind} {order} layout {layout} fp32 T=32: backward {_rel(g[b], g_ref):.2e}, "
              f"forward {_rel(co[b], c_ref):.2e}")
        assert _rel(g[b], g_ref) < 1e-6 and _rel(co[b], c_ref) < 1e-6


def test_fp32_width32_headline_shape(device, monkeypatch):
    """The headline shape (order 128, 2N-1 = 255 = 3 * 5 * 17, 136 radial nodes,
    so width-32 tiles span latitude pairs): FP32 width 32 (512 threads) is
    bit-identical to width 16 with FP32-stored grids too."""
    import torch
    import zest_amd
    from zest_amd import synth
    order, batch = 128, 2
    c = torch.from_numpy(np.stack([synth.zernike_coeffs(order, 800 + b) for b in range(batch)])).to(device)
    plan_shape = (batch,) + zest_amd.zt_shape(order)
    res = {}
    for tile in ("32", "16"):
        monkeypatch.setenv("ZGL_FFT_TILE", tile)
        plan = zest_amd.Plan(zest_amd.ZERNIKE, order, precision=zest_amd.FP32)
        g = torch.empty(plan_shape, dtype=torch.float32, device=device)
        co = torch.zeros_like(c)
        plan.backward(c, g, order, batch=batch)
        plan.forward(g, co, order, batch=batch)
        torch.cuda.synchronize()
        res[tile] = (g.cpu().numpy(), co.cpu().numpy())
    assert np.array_equal(res["32"][0], res["16"][0])
    assert np.array_equal(res["32"][1], res["16"][1])


@pytest.mark.parametrize("precision", [0, 1])
def test_wide_prime_zernike(device, monkeypatch, precision):
    """Order 192 (2N-1 = 383, a prime above 255): the matrix-core pass on eight
    waves against the VALU pass (ZGL_FFT_MFMA=0) to rounding level, FP64 and
    FP32; FP32 widths 16, 8 and 4 bit-identical."""
    from zest_amd import synth
    order = 192
    c = np.stack([synth.zernike_coeffs(order, 900 + b) for b in range(2)])
    g_m, c_m = _roundtrip(device, "zernike", order, c, precision=precision)
    monkeypatch.setenv("ZGL_FFT_MFMA", "0")
    g_v, c_v = _roundtrip(device, "zernike", order, c, precision=precision)
    tol = 1e-13 if precision == 0 else 1e-5
    for b in range(2):
        print(f"order {order} precision {precision}: grid {_rel(g_m[b], g_v[b]):.2e}, "
              f"coefficients {_rel(c_m[b], c_v[b]):.2e} vs the VALU pass")
        assert _rel(g_m[b], g_v[b]) <= tol and _rel(c_m[b], c_v[b]) <= tol
    if precision == 1:
        monkeypatch.delenv("ZGL_FFT_MFMA")
        res = {}
        for tile in ("16", "8", "4"):
            monkeypatch.setenv("ZGL_FFT_TILE", tile)
            res[tile] = _roundtrip(device, "zernike", order, c[:1], precision=precision)
        for tile in ("8", "4"):
            assert np.array_equal(res[tile][0], res["16"][0]), tile
            assert np.array_equal(res[tile][1], res["16"][1]), tile
