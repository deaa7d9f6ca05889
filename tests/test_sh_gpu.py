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


def test_strided_instances_match_packed(device):
    """Instances at non-packed strides (an odd number of doubles apart, so
    coefficient pairs are not 16-byte aligned) through the coefficient staging
    and the power spectrum: bit-identical to the packed layout."""
    import torch
    import zest_amd
    from zest_amd import synth
    L, batch = 20, 5
    c = np.stack([synth.sh_coeffs(L, 700 + b) for b in range(batch)])
    plan = zest_amd.Plan(zest_amd.SH, L)
    csz, gsz = c[0].size, int(np.prod(zest_amd.st_shape(L)))
    cs, gs = csz + 3, gsz + 5
    dc = torch.from_numpy(c).to(device)
    dg = torch.empty((batch,) + zest_amd.st_shape(L), dtype=torch.float64, device=device)
    co = torch.zeros_like(dc)
    plan.backward(dc, dg, L, batch=batch)
    plan.forward(dg, co, L, batch=batch)
    cbuf = torch.full((batch * cs,), np.nan, dtype=torch.float64, device=device)
    for b in range(batch):
        cbuf[b * cs:b * cs + csz] = dc[b].reshape(-1)
    gbuf = torch.full((batch * gs,), np.nan, dtype=torch.float64, device=device)
    obuf = torch.full((batch * cs,), np.nan, dtype=torch.float64, device=device)
    plan.backward(cbuf, gbuf, L, batch=batch, coeff_stride=cs, grid_stride=gs)
    plan.forward(gbuf, obuf, L, batch=batch, grid_stride=gs, coeff_stride=cs)
    spec = plan.power_spectrum(cbuf, L, batch=batch, coeff_stride=cs)
    torch.cuda.synchronize()
    for b in range(batch):
        assert np.array_equal(gbuf[b * gs:b * gs + gsz].cpu().numpy(), dg[b].reshape(-1).cpu().numpy())
        assert np.array_equal(obuf[b * cs:b * cs + csz].cpu().numpy(), co[b].reshape(-1).cpu().numpy())
        ref = plan.power_spectrum(dc[b:b + 1], L, batch=1)
        assert np.abs(spec[b].cpu().numpy() - ref[0].cpu().numpy()).max() <= 1e-15 * np.abs(ref.cpu().numpy()).max()


def test_sh_ragged_matches_per_order(device):
    """SH jobs of uneven orders at their own addresses (zgl_*_ragged):
    bit-identical to the per-order batched calls, and the oracle."""
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    orders = [5, 20, 12, 20, 33, 1, 12]
    plan = zest_amd.Plan(zest_amd.SH, 8)
    coeffs = [synth.sh_coeffs(N, 800 + i) for i, N in enumerate(orders)]
    jobs = [(N, torch.full(zest_amd.st_shape(N), np.nan, dtype=torch.float64, device=device),
             torch.from_numpy(c).to(device)) for N, c in zip(orders, coeffs)]
    out_jobs = [(N, g, torch.zeros_like(c)) for N, g, c in jobs]
    plan.backward_ragged(jobs)
    plan.forward_ragged(out_jobs)
    torch.cuda.synchronize()
    for i, (N, g, c) in enumerate(out_jobs):
        dc = torch.from_numpy(coeffs[i]).to(device)[None]
        dg = torch.empty((1,) + zest_amd.st_shape(N), dtype=torch.float64, device=device)
        plan.backward(dc, dg, N, batch=1)
        do = torch.zeros_like(dc)
        plan.forward(dg, do, N, batch=1)
        torch.cuda.synchronize()
        assert np.array_equal(g.cpu().numpy(), dg[0].cpu().numpy()), (i, N)
        assert np.array_equal(c.cpu().numpy(), do[0].cpu().numpy()), (i, N)
        g_ref = oracle.st_backward(coeffs[i], N)
        assert np.abs(g.cpu().numpy() - g_ref).max() <= 1e-12 * np.abs(g_ref).max()
