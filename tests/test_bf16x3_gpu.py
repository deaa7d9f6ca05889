"""The split-BF16x3 Legendre products of the FP32 modes against the oracle.

With ZGL_LEGENDRE_MATH=bf16x3 (and no autotuned choice) the Zernike
Legendre stages of the FP32 modes split each FP32 operand into three BF16
values and accumulate the six leading partial products on the BF16 MFMA in
FP32, instead of the FP32 MFMA: the error stays at the FP32 level, and one
refinement step (FP32_REFINE) still reaches FP64 accuracy. Errors are
normwise, max |x - x_ref| / max |x_ref| per instance, against the FP64 oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _run(device, order, batch, prec, c, g):
    import torch
    import zest_amd
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order, precision=prec)
    plan.autotune(False)
    dco = torch.zeros(c.shape, dtype=torch.float64, device=device)
    plan.forward(torch.from_numpy(g).to(device), dco, order, batch=batch)
    dgo = torch.empty(g.shape, dtype=torch.float64, device=device)
    plan.backward(torch.from_numpy(c).to(device), dgo, order, batch=batch)
    torch.cuda.synchronize()
    return dco.cpu().numpy(), dgo.cpu().numpy()


@pytest.mark.parametrize("order,plm", [(1, "table"), (5, "table"), (24, "table"), (64, "table"),
                                       (64, "kernel"), (128, "table"), (200, "kernel")])
def test_bf16x3_zernike(device, monkeypatch, order, plm):
    import oracle
    import zest_amd
    from zest_amd import synth
    monkeypatch.setenv("ZGL_PLM", plm)
    batch = 3
    c = np.stack([synth.zernike_coeffs(order, 40 + s) for s in range(batch)])
    g = np.stack([oracle.zt_backward(c[b], order) for b in range(batch)])
    monkeypatch.setenv("ZGL_LEGENDRE_MATH", "fp32")
    ref32 = _run(device, order, batch, zest_amd.FP32, c, g)
    monkeypatch.setenv("ZGL_LEGENDRE_MATH", "bf16x3")
    x3 = _run(device, order, batch, zest_amd.FP32, c, g)
    x3r = _run(device, order, batch, zest_amd.FP32_REFINE, c, g)
    # the BF16x3 products are a different rounding of the same sums
    assert not np.array_equal(x3[1], ref32[1]) or order <= 1
    for b in range(batch):
        c_ref = oracle.zt_forward(g[b], order)
        ef, eb = _rel(x3[0][b], c_ref), _rel(x3[1][b], g[b])
        ef32, eb32 = _rel(ref32[0][b], c_ref), _rel(ref32[1][b], g[b])
        er = _rel(x3r[0][b], c_ref)
        print(f"order {order} {plm}: bf16x3 fwd {ef:.2e} bwd {eb:.2e}; fp32 fwd {ef32:.2e} "
              f"bwd {eb32:.2e}; refined {er:.2e}")
        assert ef < 1e-6 and eb < 1e-6
        assert er < 1e-10


@pytest.mark.parametrize("tuned", [False, True])
def test_bf16x3_precision_mode(device, monkeypatch, tuned):
    """ZGL_BF16X3 plans run the BF16x3 Legendre products whatever the
    environment: the same results as FP32 plans forced by ZGL_LEGENDRE_MATH,
    and the autotuner only varies their grouping."""
    import zest_amd
    from zest_amd import synth
    import oracle
    order, batch = 48, 4
    c = np.stack([synth.zernike_coeffs(order, 70 + s) for s in range(batch)])
    g = np.stack([oracle.zt_backward(c[b], order) for b in range(batch)])
    monkeypatch.setenv("ZGL_LEGENDRE_MATH", "bf16x3")
    forced = _run(device, order, batch, zest_amd.FP32, c, g)
    monkeypatch.delenv("ZGL_LEGENDRE_MATH")
    import torch
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order, precision=zest_amd.BF16X3)
    plan.autotune(tuned)
    dco = torch.zeros(c.shape, dtype=torch.float64, device=device)
    dgo = torch.empty(g.shape, dtype=torch.float64, device=device)
    for _ in range(2):
        plan.forward(torch.from_numpy(g).to(device), dco, order, batch=batch)
        plan.backward(torch.from_numpy(c).to(device), dgo, order, batch=batch)
    torch.cuda.synchronize()
    assert np.array_equal(dco.cpu().numpy(), forced[0])
    assert np.array_equal(dgo.cpu().numpy(), forced[1])
    tuning = plan.tuning()
    if tuned:
        assert [ch for _, _, _, st, ch in tuning if st.startswith("legendre")]
        for _, rb, _, st, ch in tuning:
            assert rb == 4
            if st.startswith("legendre"):
                assert ch in (17, 18, 20, 24)
    else:
        assert tuning == []
