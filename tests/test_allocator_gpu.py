"""The device allocator hook (zgl_set_device_allocator) on the GPU.

* torch's caching allocator as the library's allocator: results bit-identical
  to hipMalloc'd scratch, the memory shows in torch's statistics while plans
  hold it and all of it comes back when they are closed.
* Out of memory injected at every allocation of plan creation and the first
  transforms in turn (zest's std::bad_alloc, include/alignment.hpp:169-176):
  each failure is ZGL_ERROR_OUT_OF_MEMORY, leaves nothing allocated once the
  plan is closed, and the first k that gets through gives the same results.
* A plan whose call failed for memory stays usable: the same call succeeds
  once memory is there, and matches the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _round_trip(plan, c, order, device):
    """backward then forward of a batch of expansions through device memory."""
    import torch
    import zest_amd
    batch = c.shape[0]
    dc = torch.from_numpy(c).to(device)
    dg = torch.empty((batch,) + plan.grid_shape(order), dtype=torch.float64, device=device)
    plan.backward(dc, dg, order, batch=batch)
    do = torch.zeros_like(dc)
    plan.forward(dg, do, order, batch=batch)
    torch.cuda.synchronize()
    assert zest_amd is not None
    return dg.cpu().numpy(), do.cpu().numpy()


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_torch_caching_allocator_as_device_allocator(device, precision):
    import torch
    import zest_amd
    from zest_amd import synth
    order, batch = 24, 5
    c = np.stack([synth.zernike_coeffs(order, 10 + s) for s in range(batch)])
    ref_plan = zest_amd.Plan(zest_amd.ZERNIKE, order, precision=precision)
    ref = _round_trip(ref_plan, c, order, device)
    ref_plan.close()
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated()
    hook = zest_amd.torch_device_allocator()
    with hook:
        plan = zest_amd.Plan(zest_amd.ZERNIKE, order, precision=precision)
        got = _round_trip(plan, c, order, device)
        assert hook.live > 0
        assert torch.cuda.memory_allocated() > before
        # a second order: tables of both are held
        live_one = hook.live
        _round_trip(plan, np.stack([synth.zernike_coeffs(17, 3)]), 17, device)
        assert hook.live > live_one
    # hipMalloc is back, but the plan's blocks still return through torch
    plan.close()
    assert hook.live == 0
    assert torch.cuda.memory_allocated() == before
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kind,precision", [(0, 0), (0, 2), (1, 1)])
def test_out_of_memory_at_every_allocation(device, kind, precision):
    """Request k + 1 fails, for k = 0, 1, ... until plan creation and a
    backward + forward pair need no more than k requests."""
    import torch
    import zest_amd
    from zest_amd import synth
    order, batch = 20, 3
    if kind == zest_amd.ZERNIKE:
        c = np.stack([synth.zernike_coeffs(order, 60 + s) for s in range(batch)])
    else:
        c = np.stack([synth.sh_coeffs(order, 60 + s) for s in range(batch)])
    ref_plan = zest_amd.Plan(kind, order, precision=precision)
    ref = _round_trip(ref_plan, c, order, device)
    ref_plan.close()
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated()
    k = 0
    while True:
        hook = zest_amd.torch_device_allocator(fail_after=k)
        plan, failure = None, None
        try:
            with hook:
                plan = zest_amd.Plan(kind, order, precision=precision)
                got = _round_trip(plan, c, order, device)
        except zest_amd.ZglError as e:
            failure = (e.status, str(e))
        # (the exception and the tensors its traceback held are gone here)
        if plan is not None:
            plan.close()
        assert hook.live == 0, (k, hook.live)
        assert torch.cuda.memory_allocated() == before, k
        if failure is None:
            break
        assert failure[0] == 2, (k, failure)
        assert "installed device allocator returned NULL" in failure[1], failure
        k += 1
        assert k < 200
    print(f"kind {kind} precision {precision}: {k} allocations")
    assert k >= 8
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_plan_usable_after_out_of_memory(device):
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 16)
    c16 = np.stack([synth.zernike_coeffs(16, 1)])
    _round_trip(plan, c16, 16, device)
    # a new order needs new tables: none can be had
    c30 = np.stack([synth.zernike_coeffs(30, 2)])
    with zest_amd.torch_device_allocator(fail_after=0):
        with pytest.raises(zest_amd.ZglError) as err:
            _round_trip(plan, c30, 30, device)
        assert err.value.status == 2
        # the order the plan holds needs nothing new and still runs
        g16, _ = _round_trip(plan, c16, 16, device)
    assert np.abs(g16[0] - oracle.zt_backward(c16[0], 16)).max() < 1e-12 * np.abs(g16[0]).max()
    # with memory available again the failed call goes through
    g30, o30 = _round_trip(plan, c30, 30, device)
    ref = oracle.zt_backward(c30[0], 30)
    assert np.abs(g30[0] - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.abs(o30[0] - c30[0]).max() < 1e-12 * np.abs(c30[0]).max()
    plan.close()
    torch.cuda.synchronize()
