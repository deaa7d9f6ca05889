"""The fused FP32 colatitude + radial kernels (zgl_fused.hip) of the Zernike
transforms, which replace the Legendre and radial stage kernels of the FP32
modes at orders up to 128.

Checked against the FP64 oracle (normwise error per instance,
max |c - c_ref| / max |c_ref|, within the 1e-6 gate of the mixed form) and
against the stage-by-stage FP32 path, over every order class
the kernels distinguish (row strips 1-4, partial radial chunks, odd and even
latitude counts, batches that leave an instance group partly empty), all
eight Zernike conventions, truncation both ways and ragged instance
addressing.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GATE = 1e-6


@pytest.fixture(autouse=True)
def _fused_on(monkeypatch):
    """The fused kernels are opt-in (ZGL_FUSED=1)."""
    monkeypatch.setenv("ZGL_FUSED", "1")


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _run(plan, c, g, order, exp_order, device):
    import torch
    batch = c.shape[0]
    dco = torch.zeros(c.shape, dtype=torch.float64, device=device)
    plan.forward(torch.from_numpy(g).to(device), dco, order, exp_order, batch=batch)
    dgo = torch.empty(g.shape, dtype=torch.float64, device=device)
    plan.backward(torch.from_numpy(c).to(device), dgo, order, exp_order, batch=batch)
    torch.cuda.synchronize()
    return dco.cpu().numpy(), dgo.cpu().numpy()


@pytest.mark.parametrize("order,batch", [(1, 1), (2, 3), (3, 9), (7, 8), (8, 2), (15, 9), (16, 5),
                                         (31, 3), (33, 10), (64, 9), (100, 3), (127, 2), (128, 9)])
def test_fused_matches_oracle_and_stagewise(device, monkeypatch, order, batch):
    import oracle
    import zest_amd
    from zest_amd import synth
    c = np.stack([synth.zernike_coeffs(order, 40 + s) for s in range(batch)])
    g = np.stack([oracle.zt_backward(c[b], order) for b in range(batch)])
    fused = zest_amd.Plan(zest_amd.ZERNIKE, order, precision=zest_amd.FP32)
    fused.profile(True)
    cf, gf = _run(fused, c, g, order, order, device)
    stats = fused.stage_stats()
    assert "legendre_radial_forward" in stats and "radial_legendre_backward" in stats, stats
    assert "legendre_forward" not in stats and "radial_backward" not in stats, stats
    monkeypatch.setenv("ZGL_FUSED", "0")
    staged = zest_amd.Plan(zest_amd.ZERNIKE, order, precision=zest_amd.FP32)
    staged.profile(True)
    cs, gs = _run(staged, c, g, order, order, device)
    assert "legendre_forward" in staged.stage_stats()
    worst = (0.0, 0.0)
    for b in range(batch):
        c_ref = oracle.zt_forward(g[b], order)
        ef, eb = _rel(cf[b], c_ref), _rel(gf[b], g[b])
        worst = (max(worst[0], ef), max(worst[1], eb))
        assert ef < GATE, (b, ef)
        assert eb < GATE, (b, eb)
        assert _rel(cf[b], cs[b]) < GATE
        assert _rel(gf[b], gs[b]) < GATE
        # the sin part at m = 0 stays exactly zero
        _, _, ms = synth.zernike_nlm(order)
        assert np.all(cf[b][ms == 0, 1] == 0.0)
    print(f"order {order} batch {batch}: fused fwd {worst[0]:.2e}, bwd {worst[1]:.2e}")


@pytest.mark.parametrize("zn,sn,sp", [(zn, sn, sp) for zn in (0, 1) for sn in (0, 1) for sp in (-1, 1)])
def test_fused_conventions(device, zn, sn, sp):
    import oracle
    import zest_amd
    from zest_amd import synth
    order, batch = 40, 3
    conv = dict(zernike_norm=zn, sh_norm=sn, sh_phase=sp)
    c = np.stack([synth.zernike_coeffs(order, 70 + s) for s in range(batch)])
    g = np.stack([oracle.zt_backward(c[b], order, **conv) for b in range(batch)])
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order, precision=zest_amd.FP32, **conv)
    cf, gf = _run(plan, c, g, order, order, device)
    for b in range(batch):
        assert _rel(cf[b], oracle.zt_forward(g[b], order, **conv)) < GATE
        assert _rel(gf[b], g[b]) < GATE


@pytest.mark.parametrize("order,exp_order", [(40, 25), (25, 40), (128, 96), (96, 128)])
def test_fused_truncation(device, order, exp_order):
    """min(exp_order, grid order) is transformed; the forward leaves the
    coefficients past it untouched and the backward reads only those below."""
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    batch = 3
    nt = min(order, exp_order)
    c = np.stack([synth.zernike_coeffs(exp_order, 90 + s) for s in range(batch)])
    g = np.stack([oracle.zt_backward(c[b], order, exp_order) for b in range(batch)])
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order, precision=zest_amd.FP32)
    sentinel = np.full(c.shape, 7.0)
    dco = torch.from_numpy(sentinel.copy()).to(device)
    plan.forward(torch.from_numpy(g).to(device), dco, order, exp_order, batch=batch)
    dgo = torch.empty(g.shape, dtype=torch.float64, device=device)
    plan.backward(torch.from_numpy(c).to(device), dgo, order, exp_order, batch=batch)
    torch.cuda.synchronize()
    ncw = zest_amd.zt_num_coeffs(nt)
    co, go = dco.cpu().numpy(), dgo.cpu().numpy()
    for b in range(batch):
        c_ref = oracle.zt_forward(g[b], order, exp_order)
        assert _rel(co[b][:ncw], c_ref[:ncw]) < GATE
        assert np.all(co[b][ncw:] == 7.0)
        assert _rel(go[b], g[b]) < GATE


def test_fused_ragged_matches_batched(device):
    """Ragged jobs address instances through offsets; their arithmetic is the
    batched one, so the results are bit-identical to the per-order calls."""
    import torch
    import zest_amd
    from zest_amd import synth
    orders = [12, 64, 12, 128, 33, 64, 12, 5, 128, 64, 33]
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 8, precision=zest_amd.FP32)
    jobs, cin = [], []
    for i, N in enumerate(orders):
        c = synth.zernike_coeffs(N, 300 + i)
        cin.append(c)
        jobs.append((N, torch.empty(zest_amd.zt_shape(N), dtype=torch.float64, device=device),
                     torch.from_numpy(c).to(device)))
    outs = [(N, g, torch.zeros_like(c)) for N, g, c in jobs]
    plan.backward_ragged(jobs)
    plan.forward_ragged(outs)
    torch.cuda.synchronize()
    for N in sorted(set(orders)):
        ids = [i for i, o in enumerate(orders) if o == N]
        dc = torch.from_numpy(np.stack([cin[i] for i in ids])).to(device)
        dg = torch.empty((len(ids),) + zest_amd.zt_shape(N), dtype=torch.float64, device=device)
        plan.backward(dc, dg, N, batch=len(ids))
        do = torch.zeros_like(dc)
        plan.forward(dg, do, N, batch=len(ids))
        torch.cuda.synchronize()
        for k, i in enumerate(ids):
            assert torch.equal(outs[i][1], dg[k]), (i, N)
            assert torch.equal(outs[i][2], do[k]), (i, N)
