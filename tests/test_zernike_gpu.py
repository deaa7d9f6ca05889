"""Zernike GLQ transforms on the GPU, checked against the oracle.

The oracle (oracle/zgl_oracle.c) restates zest's zt::GLQTransformer
(include/zernike_glq_transformer.hpp:640-949) in FP64; the tolerance for the
FP64 path is 1e-12 relative to the largest magnitude of the reference output.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CONVENTIONS = [(zn, sn, sp) for zn in (0, 1) for sn in (0, 1) for sp in (-1, 1)]
TOL = 1e-12


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _coeffs(order, seed, batch):
    from zest_amd import synth
    return np.stack([synth.zernike_coeffs(order, seed + b) for b in range(batch)])


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 16, 17, 23, 32])
def test_backward_forward_matches_oracle(device, order):
    import torch
    import oracle
    import zest_amd
    batch = 3
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order)
    c = _coeffs(order, 10, batch)
    dc = torch.from_numpy(c).to(device)
    dg = torch.empty((batch,) + zest_amd.zt_shape(order), dtype=torch.float64, device=device)
    plan.backward(dc, dg, order, batch=batch)
    # forward of a random (not band-limited) grid
    g_in = np.random.default_rng(order).standard_normal((batch,) + zest_amd.zt_shape(order))
    dgi = torch.from_numpy(g_in).to(device)
    dco = torch.zeros_like(dc)
    plan.forward(dgi, dco, order, batch=batch)
    torch.cuda.synchronize()
    g = dg.cpu().numpy()
    co = dco.cpu().numpy()
    for b in range(batch):
        assert _rel(g[b], oracle.zt_backward(c[b], order)) < TOL
        assert _rel(co[b], oracle.zt_forward(g_in[b], order)) < TOL


@pytest.mark.parametrize("conv", CONVENTIONS)
@pytest.mark.parametrize("order", [6, 11])
def test_all_conventions(device, conv, order):
    import torch
    import oracle
    import zest_amd
    zn, sn, sp = conv
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order, zernike_norm=zn, sh_norm=sn, sh_phase=sp)
    c = _coeffs(order, 3, 1)[0]
    dc = torch.from_numpy(c).to(device)
    dg = torch.empty(zest_amd.zt_shape(order), dtype=torch.float64, device=device)
    plan.backward(dc, dg, order)
    g_in = np.random.default_rng(7).standard_normal(zest_amd.zt_shape(order))
    dco = torch.zeros_like(dc)
    plan.forward(torch.from_numpy(g_in).to(device), dco, order)
    torch.cuda.synchronize()
    assert _rel(dg.cpu().numpy(), oracle.zt_backward(c, order, zernike_norm=zn, sh_norm=sn, sh_phase=sp)) < TOL
    assert _rel(dco.cpu().numpy(), oracle.zt_forward(g_in, order, zernike_norm=zn, sh_norm=sn, sh_phase=sp)) < TOL


@pytest.mark.parametrize("order", [64, 128])
def test_large_orders_match_oracle(device, order):
    import torch
    import oracle
    import zest_amd
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order)
    c = _coeffs(order, 0, 1)[0]
    dc = torch.from_numpy(c).to(device)
    dg = torch.empty(zest_amd.zt_shape(order), dtype=torch.float64, device=device)
    plan.backward(dc, dg, order)
    dco = torch.zeros_like(dc)
    plan.forward(dg, dco, order)
    torch.cuda.synchronize()
    g_ref = oracle.zt_backward(c, order)
    assert _rel(dg.cpu().numpy(), g_ref) < TOL
    assert _rel(dco.cpu().numpy(), oracle.zt_forward(g_ref, order)) < TOL
    # the quadrature is exact on band-limited grids: forward(backward(c)) = c
    assert _rel(dco.cpu().numpy(), c) < 1e-12


def test_round_trip_64_instances_order_128(device):
    """The workload of the benchmark: 64 instances at order 128, seeds 0..63."""
    import torch
    import zest_amd
    order, batch = 128, 64
    c = _coeffs(order, 0, batch)
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order)
    dc = torch.from_numpy(c).to(device)
    dg = torch.empty((batch,) + zest_amd.zt_shape(order), dtype=torch.float64, device=device)
    plan.backward(dc, dg, order, batch=batch)
    dco = torch.zeros_like(dc)
    plan.forward(dg, dco, order, batch=batch)
    torch.cuda.synchronize()
    err = (dco - dc).abs().amax(dim=(1, 2)) / dc.abs().amax(dim=(1, 2))
    assert float(err.max()) < 1e-12


@pytest.mark.parametrize("grid_order,exp_order", [(9, 5), (6, 10), (12, 1)])
def test_truncation(device, grid_order, exp_order):
    """min(exp, grid) order is used; coefficients past it are left untouched
    (include/zernike_glq_transformer.hpp:649-654, 669-675)."""
    import torch
    import oracle
    import zest_amd
    plan = zest_amd.Plan(zest_amd.ZERNIKE, grid_order)
    c = _coeffs(exp_order, 5, 1)[0]
    dc = torch.from_numpy(c).to(device)
    dg = torch.empty(zest_amd.zt_shape(grid_order), dtype=torch.float64, device=device)
    plan.backward(dc, dg, grid_order, exp_order=exp_order)
    g_in = np.random.default_rng(1).standard_normal(zest_amd.zt_shape(grid_order))
    sentinel = np.full((zest_amd.zt_num_coeffs(exp_order), 2), 7.25)
    dco = torch.from_numpy(sentinel.copy()).to(device)
    plan.forward(torch.from_numpy(g_in).to(device), dco, grid_order, exp_order=exp_order)
    torch.cuda.synchronize()
    assert _rel(dg.cpu().numpy(), oracle.zt_backward(c, grid_order, exp_order)) < TOL
    ref = oracle.zt_forward(g_in, grid_order, exp_order, out=sentinel.copy())
    assert _rel(dco.cpu().numpy(), ref) < TOL
    nt = min(grid_order, exp_order)
    assert np.all(dco.cpu().numpy()[zest_amd.zt_num_coeffs(nt):] == 7.25)


def test_host_memory_and_strides(device):
    import oracle
    import zest_amd
    order, batch = 10, 4
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order)
    nc = zest_amd.zt_num_coeffs(order)
    gsz = int(np.prod(zest_amd.zt_shape(order)))
    c = np.zeros((batch, nc + 3, 2))
    c[:, :nc] = _coeffs(order, 2, batch)
    g = np.zeros((batch, gsz + 5))
    plan.backward(c, g, order, batch=batch, coeff_stride=2 * (nc + 3), grid_stride=gsz + 5)
    co = np.zeros_like(c)
    plan.forward(g, co, order, batch=batch, grid_stride=gsz + 5, coeff_stride=2 * (nc + 3))
    for b in range(batch):
        ref = oracle.zt_backward(c[b, :nc], order).ravel()
        assert _rel(g[b, :gsz], ref) < TOL
        assert np.all(g[b, gsz:] == 0)
        assert _rel(co[b, :nc], c[b, :nc]) < 1e-12
        assert np.all(co[b, nc:] == 0)


def test_order_change_and_empty(device):
    import torch
    import oracle
    import zest_amd
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 4)
    for order in (7, 4, 7, 0):
        c = _coeffs(order, 1, 1)[0] if order else np.zeros((0, 2))
        g = np.zeros(zest_amd.zt_shape(order))
        plan.backward(c, g, order)
        assert plan.order == order
        if order:
            assert _rel(g, oracle.zt_backward(c, order)) < TOL
    # batch 0 is a no-op
    plan.forward(np.zeros(1), np.zeros(1), 5, batch=0)


@pytest.mark.parametrize("order", [9, 40])
def test_tiled_engine_matches_oracle(device, monkeypatch, order):
    """The tiled contraction engine (the fallback where the streaming kernels'
    table blocks do not fit; ZGL_LEGENDRE_ENGINE=tiled forces it) against the
    oracle, both directions."""
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    monkeypatch.setenv("ZGL_LEGENDRE_ENGINE", "tiled")
    batch = 3
    c = np.stack([synth.zernike_coeffs(order, 60 + b) for b in range(batch)])
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order)
    dc = torch.from_numpy(c).to(device)
    dg = torch.empty((batch,) + zest_amd.zt_shape(order), dtype=torch.float64, device=device)
    co = torch.zeros_like(dc)
    plan.backward(dc, dg, order, batch=batch)
    plan.forward(dg, co, order, batch=batch)
    torch.cuda.synchronize()
    g, cf = dg.cpu().numpy(), co.cpu().numpy()
    for b in range(batch):
        g_ref = oracle.zt_backward(c[b], order)
        c_ref = oracle.zt_forward(g_ref, order)
        assert np.abs(g[b] - g_ref).max() <= 1e-12 * np.abs(g_ref).max()
        assert np.abs(cf[b] - c_ref).max() <= 1e-12 * np.abs(c_ref).max()
