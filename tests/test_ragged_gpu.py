"""Ragged batches (zgl_forward_ragged / zgl_backward_ragged): instances of
uneven orders, each at its own device address, against the same transforms
run one order at a time through the batched entry points (bit-identical: the
per-instance arithmetic is the same) and against the oracle.

Covers the direct path (null stream), the captured HIP graph and its replay
(non-null stream), a changed job list, order-0 jobs, the FP32 modes and the
modelled load balance of the internal streams.
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
    assert len(costs) == 4 and min(costs) > 0
    print("lane costs (relative):", [round(c / max(costs), 3) for c in costs])


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
