"""Ragged batches (zgl_forward_ragged / zgl_backward_ragged): instances of
uneven orders, each at its own device address, against the same transforms
run one order at a time through the batched entry points (bit-identical: the
per-instance arithmetic is the same) and against the oracle.

Covers the direct path (null stream), the captured HIP graph and its replay
(non-null stream), a changed job list, order-0 jobs, the FP32 modes and the
modelled load balance. The default lanes pipeline (per-order pipelines over
four internal streams) and the persistent one (ZGL_RAGGED=persistent: one
persistent launch per stage and wave over all orders) give the same bits as
the per-order calls, the persistent one also when a small scratch budget cuts
the list into many waves and splits orders across them.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _make_jobs(device, orders, seed):
    import torch
    import zest_amd
    from zest_amd import synth
    jobs, coeffs = [], []
    for i, N in enumerate(orders):
        if N == 0:
            jobs.append((0, None, None))
            coeffs.append(None)
            continue
        c = synth.zernike_coeffs(N, seed + i)
        coeffs.append(c)
        # separate allocations: addresses in no particular order
        dc = torch.from_numpy(c).to(device)
        dg = torch.full(zest_amd.zt_shape(N), np.nan, dtype=torch.float64, device=device)
        jobs.append((N, dg, dc))
    return jobs, coeffs


def _plain(device, plan, orders, coeffs):
    """Backward then forward, one order at a time through zgl_backward/zgl_forward."""
    import torch
    import zest_amd
    grids, outs = {}, {}
    for N in sorted(set(o for o in orders if o)):
        ids = [i for i, o in enumerate(orders) if o == N]
        dc = torch.from_numpy(np.stack([coeffs[i] for i in ids])).to(device)
        dg = torch.empty((len(ids),) + zest_amd.zt_shape(N), dtype=torch.float64, device=device)
        plan.backward(dc, dg, N, batch=len(ids))
        do = torch.zeros_like(dc)
        plan.forward(dg, do, N, batch=len(ids))
        torch.cuda.synchronize()
        for k, i in enumerate(ids):
            grids[i] = dg[k].cpu().numpy()
            outs[i] = do[k].cpu().numpy()
    return grids, outs


@pytest.mark.parametrize("use_stream", [False, True])
def test_ragged_matches_per_order_and_oracle(device, use_stream):
    import torch
    import oracle
    import zest_amd
    rng = np.random.default_rng(2)
    orders = [int(x) for x in rng.integers(1, 25, size=40)] + [0, 7, 7, 24]
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 8)
    jobs, coeffs = _make_jobs(device, orders, 100)
    stream = torch.cuda.Stream() if use_stream else None
    out_jobs = [(N, g, torch.zeros_like(c) if c is not None else None) for N, g, c in jobs]
    plan.backward_ragged(jobs, stream=stream)
    plan.forward_ragged(out_jobs, stream=stream)
    torch.cuda.synchronize()
    ref_grids, ref_out = _plain(device, plan, orders, coeffs)
    for i, (N, g, c) in enumerate(out_jobs):
        if N == 0:
            continue
        assert np.array_equal(g.cpu().numpy(), ref_grids[i]), (i, N)
        assert np.array_equal(c.cpu().numpy(), ref_out[i]), (i, N)
    # oracle on a few instances
    for i in (0, 5, len(orders) - 1):
        N = orders[i]
        g_ref = oracle.zt_backward(coeffs[i], N)
        assert np.abs(ref_grids[i] - g_ref).max() <= 1e-12 * np.abs(g_ref).max()
    costs = plan.ragged_lane_costs()
    assert len(costs) >= 1 and min(costs) > 0
    print(f"{len(costs)} units, max/mean modelled work {max(costs) / (sum(costs) / len(costs)):.3f}")


def test_graph_replay_and_job_change(device):
    import torch
    import zest_amd
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 8)
    stream = torch.cuda.Stream()
    orders = [3, 9, 9, 16, 4, 12]
    jobs, coeffs = _make_jobs(device, orders, 7)
    plan.backward_ragged(jobs, stream=stream)  # capture
    torch.cuda.synchronize()
    first = [g.clone() for _, g, _ in jobs]
    for _, g, _ in jobs:
        g.fill_(np.nan)
    plan.backward_ragged(jobs, stream=stream)  # replay
    torch.cuda.synchronize()
    for a, (_, g, _) in zip(first, jobs):
        assert torch.equal(a, g)
    # new coefficient values through the same addresses: the replay reads them
    for i, (_, _, c) in enumerate(jobs):
        c.mul_(2.0)
    plan.backward_ragged(jobs, stream=stream)
    torch.cuda.synchronize()
    for a, (_, g, _) in zip(first, jobs):
        assert torch.allclose(2.0 * a, g, rtol=0, atol=1e-13 * float(a.abs().max()))
    # a different job list rebuilds the graph
    jobs2, coeffs2 = _make_jobs(device, [5, 5, 20], 50)
    plan.backward_ragged(jobs2, stream=stream)
    torch.cuda.synchronize()
    ref, _ = _plain(device, plan, [5, 5, 20], coeffs2)
    for i, (_, g, _) in enumerate(jobs2):
        assert np.array_equal(g.cpu().numpy(), ref[i])


@pytest.mark.parametrize("precision", [1, 2])
def test_ragged_fp32_modes(device, precision):
    import torch
    import oracle
    import zest_amd
    orders = [6, 11, 11, 30]
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 6, precision=precision)
    jobs, coeffs = _make_jobs(device, orders, 20)
    grids = [torch.from_numpy(oracle.zt_backward(coeffs[i], N)).to(device) for i, N in enumerate(orders)]
    out = [(N, grids[i], torch.zeros_like(jobs[i][2])) for i, N in enumerate(orders)]
    plan.forward_ragged(out, stream=torch.cuda.Stream())
    torch.cuda.synchronize()
    tol = 1e-6 if precision == 1 else 1e-10
    for i, (N, _, c) in enumerate(out):
        ref = oracle.zt_forward(grids[i].cpu().numpy(), N)
        assert np.abs(c.cpu().numpy() - ref).max() <= tol * np.abs(ref).max(), (N, precision)


def test_ragged_rejects_bad_jobs(device):
    import torch
    import zest_amd
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 4)
    c = torch.zeros((zest_amd.zt_num_coeffs(4), 2), dtype=torch.float64, device=device)
    # the binding checks the tensors against each job's order
    with pytest.raises(ValueError):
        plan.forward_ragged([(4, None, c)])
    with pytest.raises(ValueError):
        plan.forward_ragged([(1 << 20, c, c)])
    # the C ABI's own checks, reached through a raw job array
    with pytest.raises(zest_amd.ZglError):
        plan.forward_ragged(zest_amd.make_jobs([(4, None, c)]))
    with pytest.raises(zest_amd.ZglError):
        plan.forward_ragged(zest_amd.make_jobs([(1 << 20, c, c)]))
    plan.forward_ragged([])  # nothing to do


def test_ragged_job_lists_back_to_back_without_sync(device):
    """Two different job lists on one (non-blocking) stream with no host
    synchronisation between them: the second call changes the device offset
    table and drops the first call's graph only after the first call's work
    has completed, so both results are bit-identical to per-order calls."""
    import torch
    import zest_amd
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 8)
    stream = torch.cuda.Stream()
    lists = ([5, 9, 17, 9, 3, 17, 12], [20, 4, 4, 11, 20, 7])
    made = [_make_jobs(device, orders, 500 + 50 * k) for k, orders in enumerate(lists)]
    outs = [[(N, g, torch.zeros_like(c)) for N, g, c in jobs] for jobs, _ in made]
    with torch.cuda.stream(stream):
        for rep in range(2):  # the second pass replays both graphs in turn
            for (jobs, _), out in zip(made, outs):
                plan.backward_ragged(jobs, stream=stream)
                plan.forward_ragged(out, stream=stream)
    torch.cuda.synchronize()
    for (jobs, coeffs), out, orders in zip(made, outs, lists):
        ref_grids, ref_out = _plain(device, plan, orders, coeffs)
        for i, (N, g, c) in enumerate(out):
            assert np.array_equal(g.cpu().numpy(), ref_grids[i]), (i, N)
            assert np.array_equal(c.cpu().numpy(), ref_out[i]), (i, N)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("use_stream", [False, True])
def test_persistent_matches_lanes_and_waves(device, monkeypatch, precision, use_stream):
    import torch
    import zest_amd
    rng = np.random.default_rng(11)
    orders = [int(x) for x in rng.integers(8, 97, size=60)] + [96, 8, 8, 33]
    jobs, coeffs = _make_jobs(device, orders, 900)
    results = {}
    for mode, scratch_mb in (("persistent", None), ("lanes", None), ("persistent", "24"), ("persistent2", "48")):
        monkeypatch.setenv("ZGL_RAGGED", mode.rstrip("2"))
        if mode == "persistent2":  # waves alternating over two internal streams
            monkeypatch.setenv("ZGL_RAGGED_LANES", "2")
        else:
            monkeypatch.delenv("ZGL_RAGGED_LANES", raising=False)
        if scratch_mb:
            monkeypatch.setenv("ZGL_SCRATCH_MB", scratch_mb)
        else:
            monkeypatch.delenv("ZGL_SCRATCH_MB", raising=False)
        plan = zest_amd.Plan(zest_amd.ZERNIKE, 8, precision=precision)
        grids = [(N, torch.full_like(g, np.nan), c) for N, g, c in jobs]
        outs = [(N, g, torch.zeros_like(c)) for N, g, c in grids]
        stream = torch.cuda.Stream() if use_stream else None
        for _ in range(2):  # build (and capture), then run again (replay)
            plan.backward_ragged(grids, stream=stream)
            plan.forward_ragged(outs, stream=stream)
        torch.cuda.synchronize()
        results[(mode, scratch_mb)] = ([g.cpu().numpy() for _, g, _ in grids],
                                       [c.cpu().numpy() for _, _, c in outs])
        del plan
    ref = results[("lanes", None)]
    for key, (gs, cs) in results.items():
        for i in range(len(orders)):
            assert np.array_equal(gs[i], ref[0][i]), (key, i, orders[i])
            assert np.array_equal(cs[i], ref[1][i]), (key, i, orders[i])
    # and the oracle on a few instances
    import oracle
    for i in (0, len(orders) - 4):
        N = orders[i]
        g_ref = oracle.zt_backward(coeffs[i], N)
        tol = 1e-12 if precision == 0 else 1e-6
        assert np.abs(ref[0][i] - g_ref).max() <= tol * np.abs(g_ref).max()


def test_persistent_profiled_stages(device, monkeypatch):
    """With stage timing on, the persistent launches are timed per stage and
    cover every instance of the job list."""
    import zest_amd
    monkeypatch.setenv("ZGL_RAGGED", "persistent")
    orders = [20, 20, 31, 12, 64]
    jobs, _ = _make_jobs(device, orders, 70)
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 8)
    plan.profile(True)
    plan.backward_ragged(jobs)
    plan.forward_ragged([(N, g, c.clone()) for N, g, c in jobs])
    st = plan.stage_stats()
    for name in ("fft_forward", "legendre_forward", "radial_forward",
                 "radial_backward", "legendre_backward", "fft_backward"):
        assert name in st, st
        assert st[name][1] >= len(orders), (name, st[name])
