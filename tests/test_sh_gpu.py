"""Spherical-harmonic GLQ transforms on the GPU, checked against the oracle.

The oracle restates zest's st::GLQTransformer
(include/sh_glq_transformer.hpp:632-1201) in FP64, both grid layouts.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _coeffs(order, seed, batch):
    from zest_amd import synth
    return np.stack([synth.sh_coeffs(order, seed + b) for b in range(batch)])


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("order", [1, 2, 3, 5, 6, 7, 8, 13, 16, 33, 64])
def test_backward_forward_matches_oracle(device, layout, order):
    import torch
    import oracle
    import zest_amd
    batch = 5
    plan = zest_amd.Plan(zest_amd.SH, order, sh_norm=0, sh_phase=-1, layout=layout)
    c = _coeffs(order, 4, batch)
    dc = torch.from_numpy(c).to(device)
    shape = zest_amd.st_shape(order, layout)
    dg = torch.empty((batch,) + shape, dtype=torch.float64, device=device)
    plan.backward(dc, dg, order, batch=batch)
    g_in = np.random.default_rng(order).standard_normal((batch,) + shape)
    dco = torch.zeros_like(dc)
    plan.forward(torch.from_numpy(g_in).to(device), dco, order, batch=batch)
    torch.cuda.synchronize()
    for b in range(batch):
        assert _rel(dg[b].cpu().numpy(), oracle.st_backward(c[b], order, layout=layout)) < TOL
        assert _rel(dco[b].cpu().numpy(), oracle.st_forward(g_in[b], order, layout=layout)) < TOL


@pytest.mark.parametrize("conv", [(sn, sp) for sn in (0, 1) for sp in (-1, 1)])
@pytest.mark.parametrize("layout", [0, 1])
def test_all_conventions(device, conv, layout):
    import torch
    import oracle
    import zest_amd
    sn, sp = conv
    order = 9
    plan = zest_amd.Plan(zest_amd.SH, order, sh_norm=sn, sh_phase=sp, layout=layout)
    c = _coeffs(order, 8, 1)[0]
    dc = torch.from_numpy(c).to(device)
    dg = torch.empty(zest_amd.st_shape(order, layout), dtype=torch.float64, device=device)
    plan.backward(dc, dg, order)
    g_in = np.random.default_rng(2).standard_normal(zest_amd.st_shape(order, layout))
    dco = torch.zeros_like(dc)
    plan.forward(torch.from_numpy(g_in).to(device), dco, order)
    torch.cuda.synchronize()
    assert _rel(dg.cpu().numpy(), oracle.st_backward(c, order, sh_norm=sn, sh_phase=sp, layout=layout)) < TOL
    assert _rel(dco.cpu().numpy(), oracle.st_forward(g_in, order, sh_norm=sn, sh_phase=sp, layout=layout)) < TOL


def test_order_512_batch(device):
    """The precision workload's shape: L = 512, 1023 x 512 grids (a few instances)."""
    import torch
    import oracle
    import zest_amd
    order, batch = 512, 3
    plan = zest_amd.Plan(zest_amd.SH, order)
    c = _coeffs(order, 1, batch)
    dc = torch.from_numpy(c).to(device)
    dg = torch.empty((batch,) + zest_amd.st_shape(order), dtype=torch.float64, device=device)
    plan.backward(dc, dg, order, batch=batch)
    dco = torch.zeros_like(dc)
    plan.forward(dg, dco, order, batch=batch)
    torch.cuda.synchronize()
    g_ref = oracle.st_backward(c[0], order)
    assert _rel(dg[0].cpu().numpy(), g_ref) < TOL
    for b in range(batch):
        assert _rel(dco[b].cpu().numpy(), c[b]) < 1e-11


@pytest.mark.parametrize("grid_order,exp_order", [(9, 4), (5, 8)])
def test_truncation(device, grid_order, exp_order):
    import torch
    import oracle
    import zest_amd
    plan = zest_amd.Plan(zest_amd.SH, grid_order)
    c = _coeffs(exp_order, 6, 1)[0]
    dg = torch.empty(zest_amd.st_shape(grid_order), dtype=torch.float64, device=device)
    plan.backward(torch.from_numpy(c).to(device), dg, grid_order, exp_order=exp_order)
    g_in = np.random.default_rng(3).standard_normal(zest_amd.st_shape(grid_order))
    sentinel = np.full((zest_amd.st_num_coeffs(exp_order), 2), -3.5)
    dco = torch.from_numpy(sentinel.copy()).to(device)
    plan.forward(torch.from_numpy(g_in).to(device), dco, grid_order, exp_order=exp_order)
    torch.cuda.synchronize()
    assert _rel(dg.cpu().numpy(), oracle.st_backward(c, grid_order, exp_order)) < TOL
    assert _rel(dco.cpu().numpy(), oracle.st_forward(g_in, grid_order, exp_order, out=sentinel.copy())) < TOL
