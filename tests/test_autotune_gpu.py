"""The autotuner (zgl_plan_autotune): the first call at an order times the
candidate FFT tile widths and instances per workgroup (and, for the FP32
Legendre stages, FP32 against split-BF16x3 MFMA products) and keeps the
fastest. Tile widths and groupings must not change any result: each
instance's arithmetic is the same whatever the grouping, so tuned and untuned
plans agree bit for bit unless a BF16x3 product was chosen, which rounds the
same sums differently (FP32-level agreement)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_tuned_results_identical(device, precision):
    import torch
    import zest_amd
    from zest_amd import synth
    out = {}
    for tuned in (False, True):
        plan = zest_amd.Plan(zest_amd.ZERNIKE, 8, precision=precision)
        plan.autotune(tuned)
        res = []
        for order, batch in ((24, 6), (64, 5)):
            c = np.stack([synth.zernike_coeffs(order, 90 + b) for b in range(batch)])
            dc = torch.from_numpy(c).to(device)
            dg = torch.empty((batch,) + zest_amd.zt_shape(order), dtype=torch.float64, device=device)
            co = torch.zeros_like(dc)
            for _ in range(2):  # the second call runs the tuned choices
                plan.backward(dc, dg, order, batch=batch)
                plan.forward(dg, co, order, batch=batch)
            torch.cuda.synchronize()
            res.append((dg.cpu().numpy(), co.cpu().numpy()))
        out[tuned] = (res, plan.tuning())
    assert out[False][1] == []
    tuning = out[True][1]
    print("tuning:", tuning)
    x3 = any(st.startswith("legendre") and choice & 16 for _, _, _, st, choice in tuning)
    for (g0, c0), (g1, c1) in zip(out[False][0], out[True][0]):
        if x3:
            # two FP32-level results against each other, each within the 1e-6 gate
            assert np.abs(g0 - g1).max() <= 2e-6 * np.abs(g0).max()
            assert np.abs(c0 - c1).max() <= 2e-6 * np.abs(c0).max()
        else:
            assert np.array_equal(g0, g1) and np.array_equal(c0, c1)
    stages = {(o, st) for o, _, _, st, _ in tuning}
    for order in (24, 64):
        for st in ("fft_forward", "legendre_forward", "radial_forward", "radial_backward",
                   "legendre_backward", "fft_backward"):
            assert (order, st) in stages
    for o, rb, gb, st, choice in tuning:
        assert rb == (8 if precision == 0 else 4) and gb == 8
        if st.startswith("fft"):
            assert choice in (32, 16, 8, 4, 2, 1)
        elif st.startswith("legendre"):
            # instances per workgroup (+16: BF16x3 products, FP32 stages), or
            # 32 + tiles per wave (runs of equal tile counts)
            assert (choice in (1, 2, 3, 4, 6, 8, 12, 34, 35, 36, 38)
                    or (precision == 1 and choice in (17, 18, 20, 24)))
        else:
            assert choice in (1, 2, 4, 8, 16)


@pytest.mark.parametrize("precision", [0, 1])
def test_legendre_tile_runs_identical(device, monkeypatch, precision):
    """Workgroups over runs of 4 t column tiles of the flattened (instance,
    column) index (ZGL_LEGENDRE_GROUP = 32 + t) compute every output with the
    same products in the same order as instance groups: bit-identical grids
    and coefficients, at orders whose rows fill a whole number of tiles or
    not, batches smaller than one run and runs that end mid-instance."""
    import torch
    import zest_amd
    from zest_amd import synth
    for order, batch in ((128, 3), (40, 7), (9, 5), (64, 1)):
        c = np.stack([synth.zernike_coeffs(order, 30 + b) for b in range(batch)])
        dc = torch.from_numpy(c).to(device)
        res = {}
        for grp in ("", "4", "33", "35", "36", "47"):
            monkeypatch.setenv("ZGL_LEGENDRE_GROUP", grp)
            plan = zest_amd.Plan(zest_amd.ZERNIKE, order, precision=precision)
            dg = torch.empty((batch,) + zest_amd.zt_shape(order), dtype=torch.float64, device=device)
            co = torch.zeros_like(dc)
            plan.backward(dc, dg, order, batch=batch)
            plan.forward(dg, co, order, batch=batch)
            torch.cuda.synchronize()
            res[grp] = (dg.cpu().numpy(), co.cpu().numpy())
        g0, c0 = res[""]
        assert np.abs(c0 - c).max() <= (1e-10 if precision == 0 else 1e-5) * np.abs(c).max()
        for grp, (g1, c1) in res.items():
            assert np.array_equal(g0, g1), (order, batch, grp)
            assert np.array_equal(c0, c1), (order, batch, grp)


@pytest.mark.parametrize("kind,layout", [("zernike", 0), ("sh", 0), ("sh", 1)])
@pytest.mark.parametrize("precision", [0, 1])
def test_fft_tile_widths_agree(device, monkeypatch, kind, layout, precision):
    """Every FFT tile width (ZGL_FFT_TILE forces one at table build) runs the
    same per-column arithmetic: bit-identical grids and coefficients, and the
    FP64 form matches the oracle."""
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    zt = kind == "zernike"
    order, batch = (40, 3) if zt else (100, 3)
    gen = synth.zernike_coeffs if zt else synth.sh_coeffs
    c = np.stack([gen(order, 7 + b) for b in range(batch)])
    shape = zest_amd.zt_shape(order) if zt else zest_amd.st_shape(order, layout)
    res = {}
    for tile in ("32", "16", "8"):
        monkeypatch.setenv("ZGL_FFT_TILE", tile)
        plan = zest_amd.Plan(zest_amd.ZERNIKE if zt else zest_amd.SH, order, layout=layout,
                             precision=precision)
        dc = torch.from_numpy(c).to(device)
        dg = torch.empty((batch,) + shape, dtype=torch.float64, device=device)
        co = torch.zeros_like(dc)
        plan.backward(dc, dg, order, batch=batch)
        plan.forward(dg, co, order, batch=batch)
        torch.cuda.synchronize()
        res[tile] = (dg.cpu().numpy(), co.cpu().numpy())
    for tile in ("16", "8"):
        assert np.array_equal(res[tile][0], res["32"][0]) and np.array_equal(res[tile][1], res["32"][1]), tile
    if precision == 0:
        ref = oracle.zt_backward(c[0], order) if zt else oracle.st_backward(c[0], order, layout=layout)
        assert np.abs(res["32"][0][0] - ref).max() <= 1e-12 * np.abs(ref).max()
