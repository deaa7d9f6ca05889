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
        assert e32 < 1e-6
        assert eref < 1e-10
        for prec in out:
            assert _rel(out[prec][1][b], g[b]) < 1e-6


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
    for prec, tol in ((zest_amd.FP32, 1e-6), (zest_amd.FP32_REFINE, 1e-10)):
        plan = zest_amd.Plan(zest_amd.SH, order, layout=layout, precision=prec)
        dco = torch.zeros(c.shape, dtype=torch.float64, device=device)
        plan.forward(dg, dco, order, batch=batch)
        torch.cuda.synchronize()
        for b in range(batch):
            assert _rel(dco[b].cpu().numpy(), oracle.st_forward(g[b], order, layout=layout)) < tol


@pytest.mark.parametrize("order", [16, 128])
def test_zernike_grid32(device, order):
    """FP32 grids (zgl_forward_grid32 / zgl_backward_grid32). The forward
    reference is the FP64 oracle applied to the FP32 grid values widened to
    FP64, so the refined mode must reach FP64 accuracy on the grid as stored."""
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    batch = 3
    c = np.stack([synth.zernike_coeffs(order, s + 10) for s in range(batch)])
    g32 = np.stack([oracle.zt_backward(c[b], order) for b in range(batch)]).astype(np.float32)
    dg32 = torch.from_numpy(g32).to(device)
    for prec, tol in ((zest_amd.FP32, 1e-6), (zest_amd.FP32_REFINE, 1e-10)):
        plan = zest_amd.Plan(zest_amd.ZERNIKE, order, precision=prec)
        dco = torch.zeros(c.shape, dtype=torch.float64, device=device)
        plan.forward(dg32, dco, order, batch=batch)
        dgo = torch.empty_like(dg32)
        plan.backward(torch.from_numpy(c).to(device), dgo, order, batch=batch)
        torch.cuda.synchronize()
        for b in range(batch):
            c_ref = oracle.zt_forward(g32[b].astype(np.float64), order)
            e = _rel(dco[b].cpu().numpy(), c_ref)
            print(f"order {order} precision {prec}: grid32 fwd {e:.2e}")
            assert e < tol
            assert _rel(dgo[b].cpu().numpy().astype(np.float64), oracle.zt_backward(c[b], order)) < 1e-6


@pytest.mark.parametrize("layout", [0, 1])
def test_sh_grid32_host_and_device(device, layout):
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    order, batch = 48, 2
    c = np.stack([synth.sh_coeffs(order, s + 20) for s in range(batch)])
    g32 = np.stack([oracle.st_backward(c[b], order, layout=layout)
                    for b in range(batch)]).astype(np.float32)
    plan = zest_amd.Plan(zest_amd.SH, order, layout=layout, precision=zest_amd.FP32_REFINE)
    # device memory
    dco = torch.zeros(c.shape, dtype=torch.float64, device=device)
    plan.forward(torch.from_numpy(g32).to(device), dco, order, batch=batch)
    # host memory (staged, synchronous)
    hco = np.zeros_like(c)
    plan.forward(g32, hco, order, batch=batch)
    hg = np.zeros_like(g32)
    plan.backward(c, hg, order, batch=batch)
    torch.cuda.synchronize()
    for b in range(batch):
        c_ref = oracle.st_forward(g32[b].astype(np.float64), order, layout=layout)
        assert _rel(dco[b].cpu().numpy(), c_ref) < 1e-10
        assert _rel(hco[b], c_ref) < 1e-10
        assert _rel(hg[b].astype(np.float64), oracle.st_backward(c[b], order, layout=layout)) < 1e-6


def test_grid32_needs_fp32_plan(device):
    import torch
    import zest_amd
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 8, precision=zest_amd.FP64)
    g = torch.zeros(zest_amd.zt_shape(8), dtype=torch.float32, device=device)
    c = torch.zeros((zest_amd.zt_num_coeffs(8), 2), dtype=torch.float64, device=device)
    with pytest.raises(zest_amd.ZglError) as e:
        plan.forward(g, c, 8)
    assert e.value.status == 4
    with pytest.raises(zest_amd.ZglError):
        plan.backward(c, g, 8)
