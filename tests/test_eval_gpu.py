"""zgl_evaluate (zt::GridEvaluator / st::GridEvaluator) on the GPU against the
oracle, zest's own known answers (tests/test_grid_evaluator.cpp) and the GLQ
backward transform on the quadrature grid.

Tolerances: 1e-12 relative to the largest value for random expansions (FP64
throughout; the longitude table is formed directly instead of by zest's
recurrence, which differs by a few ulp times m); 1e-13 absolute for the
known answers, as zest's test uses.
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CONV_ZT = [(zn, sn, sp) for zn in (0, 1) for sn in (0, 1) for sp in (-1, 1)]


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def _points(seed, nlon, nlat, nrad):
    rng = np.random.default_rng(seed)
    lon = np.concatenate([[0.0, 2 * np.pi], rng.uniform(-2, 9, max(nlon - 2, 0))])[:nlon]
    col = np.concatenate([[0.0, np.pi], rng.uniform(0, np.pi, max(nlat - 2, 0))])[:nlat]
    r = np.concatenate([[0.0, 1.0], rng.uniform(0, 1.05, max(nrad - 2, 0))])[:nrad]
    return lon, col, r


def _dev(device, *arrays):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(device) for a in arrays]


@pytest.mark.parametrize("order,nlon,nlat,nrad,batch", [
    (1, 3, 2, 1, 1), (2, 5, 4, 3, 2), (7, 17, 9, 11, 3), (16, 33, 20, 9, 2),
    (33, 70, 41, 37, 2), (64, 129, 65, 70, 3), (130, 40, 30, 20, 1)])
def test_zernike_matches_oracle(device, order, nlon, nlat, nrad, batch):
    import oracle
    import zest_amd
    from zest_amd import synth
    print(f"zernike evaluate order={order} grid={nlon}x{nlat}x{nrad} batch={batch}", flush=True)
    lon, col, r = _points(order, nlon, nlat, nrad)
    cs = [synth.zernike_coeffs(order, 100 + i) for i in range(batch)]
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 4)
    dc, dl, dt, dr = _dev(device, np.stack(cs), lon, col, r)
    v = plan.evaluate(dc, order, dl, dt, dr, batch=batch).cpu().numpy()
    for b in range(batch):
        ref = oracle.zt_evaluate(cs[b], order, lon, col, r)
        assert rel(v[b], ref) < 1e-12, (order, b)


@pytest.mark.parametrize("conv", CONV_ZT)
def test_zernike_conventions(device, conv):
    import oracle
    import zest_amd
    from zest_amd import synth
    zn, sn, sp = conv
    order = 12
    lon, col, r = _points(5, 23, 14, 9)
    c = synth.zernike_coeffs(order, 9)
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order, zernike_norm=zn, sh_norm=sn, sh_phase=sp)
    dc, dl, dt, dr = _dev(device, c, lon, col, r)
    v = plan.evaluate(dc, order, dl, dt, dr).cpu().numpy()[0]
    assert rel(v, oracle.zt_evaluate(c, order, lon, col, r, zn, sn, sp)) < 1e-12


@pytest.mark.parametrize("order,nlon,nlat,batch", [
    (1, 4, 3, 1), (5, 11, 6, 5), (17, 35, 18, 9), (100, 199, 101, 3), (24, 50, 25, 70)])
@pytest.mark.parametrize("conv", [(0, -1), (1, 1)])
def test_sh_matches_oracle(device, order, nlon, nlat, batch, conv):
    import oracle
    import zest_amd
    from zest_amd import synth
    sn, sp = conv
    print(f"sh evaluate order={order} grid={nlon}x{nlat} batch={batch}", flush=True)
    lon, col, _ = _points(order + 1, nlon, nlat, 0)
    cs = [synth.sh_coeffs(order, 200 + i) for i in range(batch)]
    plan = zest_amd.Plan(zest_amd.SH, 4, sh_norm=sn, sh_phase=sp, layout=1)
    dc, dl, dt = _dev(device, np.stack(cs), lon, col)
    v = plan.evaluate(dc, order, dl, dt, batch=batch).cpu().numpy()
    assert v.shape == (batch, nlon, nlat)
    for b in range(batch):
        assert rel(v[b], oracle.st_evaluate(cs[b], order, lon, col, sn, sp)) < 1e-12, b


def _linspace(a, b, n):
    x = np.linspace(a, b, n)
    x[-1] = b
    return x


@pytest.mark.parametrize("conv", CONV_ZT)
def test_known_answers(device, conv):
    """tests/test_grid_evaluator.cpp:57-415, through the GPU path."""
    import zest_amd
    zn, sn, sp = conv
    order = 6
    shnorm = 0.5 / math.sqrt(math.pi) if sn == 1 else 1.0
    phase = -1.0 if sp == -1 else 1.0
    # SH: Y_31 + Y_4,-3 on 30 x 15
    lon, col = _linspace(0, 2 * np.pi, 30), _linspace(0, np.pi, 15)
    L, C = np.meshgrid(lon, col, indexing="ij")
    z = np.cos(C)
    c = np.zeros((zest_amd.st_num_coeffs(order), 2))
    c[3 * 4 // 2 + 1, 0] = 1.0
    c[4 * 5 // 2 + 3, 1] = 1.0
    f = phase * shnorm * (math.sqrt(21 / 8) * np.sqrt(1 - z * z) * (5 * z * z - 1) * np.cos(L)
                          + math.sqrt(315 / 8) * np.sqrt(1 - z * z) * (1 - z * z) * z * np.sin(3 * L))
    plan = zest_amd.Plan(zest_amd.SH, order, sh_norm=sn, sh_phase=sp)
    dc, dl, dt = _dev(device, c, lon, col)
    assert np.abs(plan.evaluate(dc, order, dl, dt).cpu().numpy()[0] - f).max() < 1e-13
    # Zernike: Z_33,-2 + Z_531 on 14 x 7 x 7
    lon, col, r = _linspace(0, 2 * np.pi, 14), _linspace(0, np.pi, 7), _linspace(0, 1, 7)
    L, C, R = np.meshgrid(lon, col, r, indexing="ij")
    z = np.cos(C)
    tet = lambda n, l, m: (n + 1) * (n + 3) * (2 * n + 1) // 24 + l * l // 4 + m
    c = np.zeros((zest_amd.zt_num_coeffs(order), 2))
    c[tet(3, 3, 2), 1] = 1.0
    c[tet(5, 3, 1), 0] = 1.0
    zn3 = 3.0 if zn == 0 else 1.0
    zn5 = math.sqrt(13) if zn == 0 else 1.0
    f = shnorm * (zn3 * R ** 3 * math.sqrt(105 / 4) * (1 - z * z) * z * np.sin(2 * L)
                  + zn5 * phase * (5.5 * R * R - 4.5) * R ** 3 * math.sqrt(21 / 8)
                  * np.sqrt(1 - z * z) * (5 * z * z - 1) * np.cos(L))
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order, zernike_norm=zn, sh_norm=sn, sh_phase=sp)
    dc, dl, dt, dr = _dev(device, c, lon, col, r)
    assert np.abs(plan.evaluate(dc, order, dl, dt, dr).cpu().numpy()[0] - f).max() < 1e-13


@pytest.mark.parametrize("order", [16, 64])
def test_equals_backward_on_quadrature_grid(device, order):
    """On the quadrature grid the direct sums equal the GLQ backward transform."""
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    nphi, nth, nr = zest_amd.zt_shape(order)
    lon = 2 * np.pi * np.arange(nphi) / nphi
    z, _ = oracle.gl_nodes_weights(nth)
    x, _ = oracle.gl_nodes_weights(nr)
    r = 0.5 * (1 + x)
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order)
    c = synth.zernike_coeffs(order, 3)
    dc, dl, dt, dr = _dev(device, c, lon, np.arccos(z), r)
    v = plan.evaluate(dc, order, dl, dt, dr)
    g = torch.empty((1,) + zest_amd.zt_shape(order), dtype=torch.float64, device=device)
    plan.backward(dc, g, order)
    torch.cuda.synchronize()
    assert rel(v.cpu().numpy(), g.cpu().numpy()) < 1e-12


def test_host_memory_strides_and_empty(device):
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    order, batch = 9, 3
    lon, col, r = _points(1, 13, 8, 6)
    cs = np.stack([synth.zernike_coeffs(order, i) for i in range(batch)])
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order)
    vh = plan.evaluate(cs, order, lon, col, r, batch=batch)  # host arrays
    dc, dl, dt, dr = _dev(device, cs, lon, col, r)
    vd = plan.evaluate(dc, order, dl, dt, dr, batch=batch).cpu().numpy()
    assert np.array_equal(vh, vd)
    # padded output stride: the gaps stay untouched
    npts = 13 * 8 * 6
    out = torch.full((batch, npts + 5), 7.0, dtype=torch.float64, device=device)
    plan.evaluate(dc, order, dl, dt, dr, values=out, batch=batch, value_stride=npts + 5)
    o = out.cpu().numpy()
    assert np.array_equal(o[:, :npts].reshape(vd.shape), vd)
    assert (o[:, npts:] == 7.0).all()
    for b in range(batch):
        assert rel(vd[b], oracle.zt_evaluate(cs[b], order, lon, col, r)) < 1e-12
    # order 0 writes zeros; an empty axis writes nothing
    z0 = plan.evaluate(dc, 0, dl, dt, dr, values=torch.full_like(out[:, :npts], 3.0), batch=batch)
    assert (z0.cpu().numpy() == 0).all()
    e = plan.evaluate(dc, order, dl[:0], dt, dr)
    assert e.shape == (1, 0, 8, 6)


def test_plm_in_kernel_matches_table(device, monkeypatch):
    import oracle
    import zest_amd
    from zest_amd import synth
    order = 40
    lon, col, r = _points(2, 31, 45, 10)
    c = synth.zernike_coeffs(order, 4)
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 4)
    dc, dl, dt, dr = _dev(device, c, lon, col, r)
    monkeypatch.setenv("ZGL_PLM", "table")
    a = plan.evaluate(dc, order, dl, dt, dr).cpu().numpy()
    monkeypatch.setenv("ZGL_PLM", "kernel")
    b = plan.evaluate(dc, order, dl, dt, dr).cpu().numpy()
    assert rel(b, a) < 1e-13
    assert rel(b[0], oracle.zt_evaluate(c, order, lon, col, r)) < 1e-12


@pytest.mark.parametrize("plm", ["table", "kernel"])
def test_order_200(device, monkeypatch, plm):
    """Order 200: the tiled radial kernel (beyond the streaming one's 64 n per
    l), two or one 16-point strips in the Legendre stage, both P sources."""
    import oracle
    import zest_amd
    from zest_amd import synth
    order = 200
    monkeypatch.setenv("ZGL_PLM", plm)
    lon, col, r = _points(8, 21, 17, 12)
    c = synth.zernike_coeffs(order, 1)
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 4)
    dc, dl, dt, dr = _dev(device, c, lon, col, r)
    v = plan.evaluate(dc, order, dl, dt, dr).cpu().numpy()[0]
    assert rel(v, oracle.zt_evaluate(c, order, lon, col, r)) < 1e-12


def test_sh_order_700_one_strip(device):
    """L = 700: the Legendre stage's P block fits only as one 16-point strip."""
    import oracle
    import zest_amd
    from zest_amd import synth
    order = 700
    lon, col, _ = _points(9, 19, 40, 0)
    cs = np.stack([synth.sh_coeffs(order, i) for i in range(3)])
    plan = zest_amd.Plan(zest_amd.SH, 4)
    dc, dl, dt = _dev(device, cs, lon, col)
    v = plan.evaluate(dc, order, dl, dt, batch=3).cpu().numpy()
    for b in range(3):
        assert rel(v[b], oracle.st_evaluate(cs[b], order, lon, col)) < 1e-12
