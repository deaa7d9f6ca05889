"""The benchmark configurations run end to end on the GPU, against zest's
golden fixtures (tests/golden/zest_golden.npz, written by compiled zest) and
the oracle:

  N=32, seed 0          the reference's CPU-runnable case
  N=128, 64 instances   seeds 0..63; instances 0-3 against zest's fixtures
  SH L=512              FP64, FP32 and FP32 + refinement against FP64 (seed 1
                        for instance 0: zest's fixture)
  10,000 ragged         orders U{8..96} (seed 2): a subset bit-identical to
                        per-order calls and matching the oracle, the whole set
                        through the round trip
  N=64 fits             FP32 gradients against FP64 central differences

Errors: normwise per instance, max|x - x_ref| / max|x_ref| (the 1e-6 gate of
the mixed forms), and, reported beside it, elementwise with a floor:
max |x - x_ref| / max(|x_ref|, 1e-3 max|x_ref|).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "zest_golden.npz")
GATE = 1e-6


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _elem(a, b, floor=1e-3):
    b = np.asarray(b)
    return float((np.abs(a - b) / np.maximum(np.abs(b), floor * np.abs(b).max())).max())


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


def test_zt32_seed0_against_zest(device, golden):
    import torch
    import zest_amd
    c = golden["zt32_coeffs"]
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 32)
    dg = torch.empty(zest_amd.zt_shape(32), dtype=torch.float64, device=device)
    plan.backward(torch.from_numpy(c).to(device), dg, 32)
    dc = torch.zeros_like(torch.from_numpy(c)).to(device)
    plan.forward(dg, dc, 32)
    torch.cuda.synchronize()
    g = dg.cpu().numpy()
    vals = golden["zt32_grid_val"]
    assert _rel(g.ravel()[golden["zt32_grid_idx"]], vals) < 1e-12
    sums = golden["zt32_grid_sums"]
    assert abs(g.sum() - sums[0]) <= 1e-12 * np.abs(g).sum()
    assert abs((g * g).sum() - sums[1]) <= 1e-12 * sums[1]
    # zest's forward of its own backward grid; the forward recovers the seeded coefficients
    assert _rel(dc.cpu().numpy(), golden["zt32_fwd"]) < 1e-12
    assert _rel(dc.cpu().numpy(), c) < 1e-12


def test_zt128_batch64_against_zest(device, golden):
    """The FP64 drop-in workload: 64 instances at order 128, seeds 0..63, in
    one batched call each way; instances 0-3 against zest's fixtures, every
    instance through the round trip; the mixed forms (FP32 arithmetic, FP64 and
    FP32 grids) of instance 0 against the same fixtures at the 1e-6 gate."""
    import torch
    import zest_amd
    from zest_amd import synth
    N, B = 128, 64
    c = np.stack([synth.zernike_coeffs(N, s) for s in range(B)])
    dc = torch.from_numpy(c).to(device)
    plan = zest_amd.Plan(zest_amd.ZERNIKE, N)
    dg = torch.empty((B,) + zest_amd.zt_shape(N), dtype=torch.float64, device=device)
    plan.backward(dc, dg, N, batch=B)
    dco = torch.zeros_like(dc)
    plan.forward(dg, dco, N, batch=B)
    torch.cuda.synchronize()
    co = dco.cpu().numpy()
    for b in range(B):
        assert _rel(co[b], c[b]) < 1e-12, b
    g4 = dg[:4].cpu().numpy()
    for s in range(4):
        g = g4[s]
        assert _rel(g.ravel()[golden[f"zt128_{s}_grid_idx"]], golden[f"zt128_{s}_grid_val"]) < 1e-12
        sums = golden[f"zt128_{s}_grid_sums"]
        assert abs((g * g).sum() - sums[1]) <= 1e-12 * sums[1]
        cidx = golden[f"zt128_{s}_fwd_idx"]
        assert _rel(co[s][cidx], golden[f"zt128_{s}_fwd_val"]) < 1e-12
    p32 = zest_amd.Plan(zest_amd.ZERNIKE, N, precision=zest_amd.FP32)
    for gdt in (torch.float64, torch.float32):
        g32 = torch.empty((1,) + zest_amd.zt_shape(N), dtype=gdt, device=device)
        p32.backward(dc[:1], g32, N)
        c32 = torch.zeros_like(dc[:1])
        p32.forward(dg[:1].to(gdt), c32, N)
        torch.cuda.synchronize()
        gv = g32[0].double().cpu().numpy().ravel()[golden["zt128_0_grid_idx"]]
        eb = _rel(gv, golden["zt128_0_grid_val"])
        cidx = golden["zt128_0_fwd_idx"]
        ef = _rel(c32[0].cpu().numpy()[cidx], golden["zt128_0_fwd_val"])
        print(f"order 128 FP32 ({gdt}): backward {eb:.2e}, forward {ef:.2e} vs zest")
        assert eb < GATE and ef < GATE


def test_sh512_precision_against_fp64(device, golden):
    """SH at L=512 (LonLat grids 1023 x 512, conventions geo / none), four
    instances, seeds 1..4 (N(0,1)/(1+l)^1.5 coefficients from the portable
    generator; seed 1 is zest's fixture): the FP64 transforms against zest's
    fixture and the oracle, the FP32 forward within the 1e-6 gate, the FP32 +
    refinement forward at FP64 level."""
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    L, B = 512, 4
    c = np.stack([synth.sh_coeffs(L, 1 + b) for b in range(B)])
    dc = torch.from_numpy(c).to(device)
    p64 = zest_amd.Plan(zest_amd.SH, L)
    dg = torch.empty((B,) + zest_amd.st_shape(L), dtype=torch.float64, device=device)
    p64.backward(dc, dg, L, batch=B)
    c64 = torch.zeros_like(dc)
    p64.forward(dg, c64, L, batch=B)
    out = {}
    for name, prec in (("fp32", zest_amd.FP32), ("fp32+refine", zest_amd.FP32_REFINE)):
        plan = zest_amd.Plan(zest_amd.SH, L, precision=prec)
        cx = torch.zeros_like(dc)
        plan.forward(dg, cx, L, batch=B)
        gx = torch.empty_like(dg)
        plan.backward(dc, gx, L, batch=B)
        out[name] = (cx, gx)
    torch.cuda.synchronize()
    g = dg.cpu().numpy()
    f64 = c64.cpu().numpy()
    assert _rel(g[0].ravel()[golden["st512_grid_idx"]], golden["st512_grid_val"]) < 1e-12
    assert abs((g[0] * g[0]).sum() - golden["st512_grid_sums"][1]) <= 1e-12 * golden["st512_grid_sums"][1]
    assert _rel(f64[0][golden["st512_fwd_idx"]], golden["st512_fwd_val"]) < 1e-12
    g_ref = oracle.st_backward(c[B - 1], L)
    assert _rel(g[B - 1], g_ref) < 1e-12
    assert _rel(f64[B - 1], oracle.st_forward(g_ref, L)) < 1e-12
    for name, (cx, gx) in out.items():
        cxh, gxh = cx.cpu().numpy(), gx.cpu().numpy()
        ef = max(_rel(cxh[b], f64[b]) for b in range(B))
        ee = max(_elem(cxh[b], f64[b]) for b in range(B))
        eb = max(_rel(gxh[b], g[b]) for b in range(B))
        print(f"SH L=512 {name}: forward {ef:.2e} normwise, {ee:.2e} elementwise (floor 1e-3); "
              f"backward {eb:.2e}")
        assert eb < GATE
        assert ef < (1e-10 if name == "fp32+refine" else GATE)


def _ragged_orders():
    from zest_amd import synth
    return (8 + np.floor(synth.uniform(2, 10000) * 89)).astype(int)


def test_ragged_10000_u8_96(device):
    """10,000 Zernike instances of orders U{8..96} (seed 2), each at its own
    address: the first 300 (instance order) bit-identical to per-order
    batched calls on the same inputs, five of them against the oracle, and all
    10,000 through the ragged round trip forward(backward(c)) = c."""
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    orders = _ragged_orders()
    assert orders.min() >= 8 and orders.max() <= 96 and len(set(orders.tolist())) == 89
    gen = torch.Generator(device=device)
    gen.manual_seed(2)
    coeffs, grids, outs = [None] * len(orders), [None] * len(orders), [None] * len(orders)
    groups = {}
    for o in sorted(set(orders.tolist())):
        ids = np.nonzero(orders == o)[0]
        n, _, m = synth.zernike_nlm(o)
        c = torch.randn((len(ids), len(n), 2), generator=gen, dtype=torch.float64, device=device)
        c *= torch.from_numpy(1.0 / (1.0 + n)).to(device)[None, :, None]
        c[:, torch.from_numpy(m == 0).to(device), 1] = 0.0
        g = torch.empty((len(ids),) + zest_amd.zt_shape(o), dtype=torch.float64, device=device)
        co = torch.zeros_like(c)
        groups[o] = (c, co)
        for k, i in enumerate(ids):
            coeffs[i], grids[i], outs[i] = c[k], g[k], co[k]
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 8)
    stream = torch.cuda.Stream()
    jb = zest_amd.make_jobs([(int(o), grids[i], coeffs[i]) for i, o in enumerate(orders)])
    jf = zest_amd.make_jobs([(int(o), grids[i], outs[i]) for i, o in enumerate(orders)])
    plan.backward_ragged(jb, stream=stream)
    plan.forward_ragged(jf, stream=stream)
    torch.cuda.synchronize()
    worst = 0.0
    for c, co in groups.values():
        d = (co - c).abs().amax(dim=(1, 2)) / c.abs().amax(dim=(1, 2))
        worst = max(worst, float(d.max()))
    print(f"10,000 ragged instances: worst round-trip error {worst:.2e}")
    assert worst < 1e-12
    # the first 300 against per-order batched calls on copies of the same inputs
    sub = list(range(300))
    for o in sorted(set(orders[sub].tolist())):
        ids = [i for i in sub if orders[i] == o]
        dc = torch.stack([coeffs[i] for i in ids])
        dg = torch.empty((len(ids),) + zest_amd.zt_shape(o), dtype=torch.float64, device=device)
        plan.backward(dc, dg, o, batch=len(ids))
        do = torch.zeros_like(dc)
        plan.forward(dg, do, o, batch=len(ids))
        torch.cuda.synchronize()
        for k, i in enumerate(ids):
            assert torch.equal(grids[i], dg[k]), (i, o)
            assert torch.equal(outs[i], do[k]), (i, o)
    for i in (0, 1, 57, 150, 299):
        o = int(orders[i])
        g_ref = oracle.zt_backward(coeffs[i].cpu().numpy(), o)
        assert _rel(grids[i].cpu().numpy(), g_ref) < 1e-12, (i, o)
        assert _rel(outs[i].cpu().numpy(), oracle.zt_forward(g_ref, o)) < 1e-12, (i, o)


def test_fp32_gradient_against_fp64_differences(device):
    """Least-squares fits at order 64 (targets r exp(-(a.x)^2), a ~ U(0.5, 2),
    seed 3): the gradient of sum_x W_x (B c - g)^2 from the FP32 plan's
    hand-written adjoint, against central differences of the FP64 loss
    (h = 1e-6; the loss is quadratic, so they are exact up to rounding) on
    sampled components, and against the FP64 adjoint's gradient."""
    import torch
    import zest_amd
    from zest_amd import synth
    from zest_amd.autograd import ZernikeBackward
    from test_adjoint_gpu import _fit_problem
    order, nfit = 64, 4
    g, W, _ = _fit_problem(order, 3, nfit)
    c0 = np.stack([synth.zernike_coeffs(order, 300 + i) for i in range(nfit)]) * 0.1
    dW, dg = torch.from_numpy(W).to(device), torch.from_numpy(g).to(device)
    grads = {}
    plans = {p: zest_amd.Plan(zest_amd.ZERNIKE, order, precision=p) for p in (zest_amd.FP64, zest_amd.FP32)}
    for p, plan in plans.items():
        def loss(c, plan=plan):
            r = ZernikeBackward.apply(c, plan, order) - dg
            return (dW * r * r).sum(dim=(1, 2, 3))
        c = torch.from_numpy(c0).to(device).requires_grad_(True)
        loss(c).sum().backward()
        grads[p] = c.grad.detach().cpu().numpy()
    g64, g32 = grads[zest_amd.FP64], grads[zest_amd.FP32]
    e = max(_rel(g32[i], g64[i]) for i in range(nfit))
    print(f"FP32 gradient vs FP64 gradient: {e:.2e} normwise")
    assert e < GATE

    def loss64(c):
        r = ZernikeBackward.apply(c, plans[zest_amd.FP64], order) - dg
        return (dW * r * r).sum(dim=(1, 2, 3))
    m = synth.zernike_nlm(order)[2]
    rng = np.random.default_rng(6)
    h = 1e-6
    c = torch.from_numpy(c0).to(device)
    worst = 0.0
    with torch.no_grad():
        for k in rng.choice(g32.shape[1], 16, replace=False):
            for part in (0, 1):
                if part == 1 and m[k] == 0:
                    continue
                cp, cm = c.clone(), c.clone()
                cp[:, k, part] += h
                cm[:, k, part] -= h
                fd = ((loss64(cp) - loss64(cm)) / (2 * h)).cpu().numpy()
                for i in range(nfit):
                    worst = max(worst, abs(fd[i] - g32[i, k, part]) / np.abs(g64[i]).max())
    print(f"FP32 gradient vs FP64 central differences: {worst:.2e} (relative to max |gradient|)")
    assert worst < GATE
