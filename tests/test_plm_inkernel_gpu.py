"""P_lm blocks recomputed inside the Legendre kernels instead of read from a
stored table (the default for Zernike orders >= 192; ZGL_PLM=kernel / table
forces either at plan creation). Both modes run the same recursion
(zest_amd/csrc/zgl_plm.hpp); results must agree to rounding, match the
oracle / compiled zest, and keep the round trip forward(backward(c)) = c."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plan(mode, order, precision=0):
    import zest_amd
    old = os.environ.get("ZGL_PLM")
    os.environ["ZGL_PLM"] = mode
    try:
        return zest_amd.Plan(zest_amd.ZERNIKE, order, precision=precision)
    finally:
        if old is None:
            del os.environ["ZGL_PLM"]
        else:
            os.environ["ZGL_PLM"] = old


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _run(plan, order, c, device):
    import torch
    import zest_amd
    dc = torch.from_numpy(c).to(device)
    batch = c.shape[0]
    dg = torch.empty((batch,) + zest_amd.zt_shape(order), dtype=torch.float64, device=device)
    plan.backward(dc, dg, order, batch=batch)
    co = torch.zeros_like(dc)
    plan.forward(dg, co, order, batch=batch)
    torch.cuda.synchronize()
    return dg.cpu().numpy(), co.cpu().numpy()


@pytest.mark.parametrize("order", [5, 16, 64, 128])
@pytest.mark.parametrize("precision,tol", [(0, 1e-13), (1, 1e-6)])
def test_kernel_and_table_agree(device, order, precision, tol):
    import oracle
    from zest_amd import synth
    c = np.stack([synth.zernike_coeffs(order, 60 + b) for b in range(2)])
    gk, ck = _run(_plan("kernel", order, precision), order, c, device)
    gt, ct = _run(_plan("table", order, precision), order, c, device)
    assert _rel(gk, gt) < tol and _rel(ck, ct) < tol
    if precision == 0 and order <= 64:
        assert _rel(gk[0], oracle.zt_backward(c[0], order)) < 1e-12
        assert _rel(ck[0], oracle.zt_forward(gk[0], order)) < 1e-12


def test_order_256_recomputes_plm(device):
    import oracle
    from zest_amd import synth
    order = 256
    c = synth.zernike_coeffs(order, 3)[None]
    g, co = _run(_plan("auto", order), order, c, device)
    # forward(backward(c)) = c: the quadrature is exact for band-limited input
    assert _rel(co[0], c[0]) < 1e-12
    gt, ct = _run(_plan("table", order), order, c, device)
    assert _rel(g, gt) < 1e-13 and _rel(co, ct) < 1e-13
    if oracle.REF is not None:
        z = oracle.ZestZernike(order, 1, 0, -1)
        g_ref = z.backward(c[0])
        assert _rel(g[0], g_ref) < 1e-12
        assert _rel(co[0], z.forward(g[0])) < 1e-12
