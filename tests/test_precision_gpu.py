"""FP32 and FP32 + FP64-residual refinement modes against the FP64 oracle.

The mixed forms run the FFT and the Legendre/radial MFMA contractions in FP32
(FP32 accumulation); the tables are generated in FP64 and rounded. Errors are
normwise: max |c - c_ref| / max |c_ref| per instance.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("order", [8, 32, 128])
def test_zernike_fp32_and_refined(device, order):
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    batch = 2
    c = np.stack([synth.zernike_coeffs(order, s) for s in range(batch)])
    g = np.stack([oracle.zt_backward(c[b], order) for b in range(batch)])
    dg = torch.from_numpy(g).to(device)
    out = {}
    for prec in (zest_amd.FP32, zest_amd.FP32_REFINE):
        plan = zest_amd.Plan(zest_amd.ZERNIKE, order, precision=prec)
        dco = torch.zeros(c.shape, dtype=torch.float64, device=device)
        plan.forward(dg, dco, order, batch=batch)
        dgo = torch.empty_like(dg)
        plan.backward(torch.from_numpy(c).to(device), dgo, order, batch=batch)
        torch.cuda.synchronize()
        out[prec] = (dco.cpu().numpy(), dgo.cpu().numpy())
    for b in range(batch):
        c_ref = oracle.zt_forward(g[b], order)
        e32 = _rel(out[zest_amd.FP32][0][b], c_ref)
        eref = _rel(out[zest_amd.FP32_REFINE][0][b], c_ref)
        print(f"order {order}: fp32 fwd {e32:.2e}, refined fwd {eref:.2e}")
        assert e32 < 5e-6
        assert eref < 1e-10
        for prec in out:
            assert _rel(out[prec][1][b], g[b]) < 5e-6


@pytest.mark.parametrize("layout", [0, 1])
def test_sh_fp32_and_refined(device, layout):
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    order, batch = 64, 3
    c = np.stack([synth.sh_coeffs(order, s) for s in range(batch)])
    g = np.stack([oracle.st_backward(c[b], order, layout=layout) for b in range(batch)])
    dg = torch.from_numpy(g).to(device)
    for prec, tol in ((zest_amd.FP32, 5e-6), (zest_amd.FP32_REFINE, 1e-10)):
        plan = zest_amd.Plan(zest_amd.SH, order, layout=layout, precision=prec)
        dco = torch.zeros(c.shape, dtype=torch.float64, device=device)
        plan.forward(dg, dco, order, batch=batch)
        torch.cuda.synchronize()
        for b in range(batch):
            assert _rel(dco[b].cpu().numpy(), oracle.st_forward(g[b], order, layout=layout)) < tol
