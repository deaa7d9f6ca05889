"""zgl_forward_noalloc, the forward transform without any allocation (scratch
from the library's static device array), against zgl_forward on the same
inputs: bit-identical, since both run the same kernels with the same tile
shapes and a column's arithmetic does not depend on how many instances share
a chunk. Two calls in a row, the second with a smaller input (lower order,
fewer instances), for each precision and both transform kinds, one of them
split into several chunks by the size of the static array; and the two
requirements the entry point reports as errors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(plan, grid, coeffs, grid_order, batch, stream):
    """Through the C ABI directly."""
    import zest_amd
    lib = zest_amd.LIB
    return lib.zgl_forward_noalloc(plan._h, grid.data_ptr(), grid_order, coeffs.data_ptr(), grid_order,
                                   batch, 0, 0, stream.cuda_stream)


def _inputs(device, plan, kind, order, batch, seed):
    import torch
    import zest_amd
    from zest_amd import synth
    gen = synth.zernike_coeffs if kind == zest_amd.ZERNIKE else synth.sh_coeffs
    c = torch.from_numpy(np.stack([gen(order, seed + b) for b in range(batch)])).to(device)
    g = torch.empty((batch,) + plan.grid_shape(order), dtype=torch.float64, device=device)
    plan.backward(c, g, order, batch=batch)
    return c, g


CASES = [
    # kind, precision, (order, batch) of the first and of the second call
    ("zernike", 0, (40, 7), (24, 3)),
    ("zernike", 1, (64, 9), (33, 2)),
    ("zernike", 2, (48, 5), (20, 1)),
    ("zernike", 0, (128, 5), (96, 3)),   # 2 (then 4) instances per chunk of the static array
    ("sh", 0, (200, 6), (120, 2)),
    ("sh", 1, (64, 4), (31, 3)),
]


@pytest.mark.parametrize("kind,precision,first,second", CASES)
def test_noalloc_matches_forward(device, kind, precision, first, second):
    import torch
    import zest_amd
    k = zest_amd.ZERNIKE if kind == "zernike" else zest_amd.SH
    plan = zest_amd.Plan(k, first[0], precision=precision)
    plan.resize(second[0])  # the tables of both orders exist before the calls
    stream = torch.cuda.current_stream()
    for call, (order, batch) in enumerate((first, second)):
        print(f"{kind} precision {precision} call {call}: order {order}, {batch} instances", flush=True)
        c, g = _inputs(device, plan, k, order, batch, 100 * call + 7)
        ref = torch.zeros_like(c)
        plan.forward(g, ref, order, batch=batch)
        out = torch.zeros_like(c)
        assert _noalloc(plan, g, out, order, batch, stream) == 0, zest_amd.LIB.zgl_last_error_string()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (kind, precision, call)
        assert plan.order == order


def test_noalloc_requirements(device):
    import torch
    import zest_amd
    stream = torch.cuda.current_stream()
    lib = zest_amd.LIB
    # the tables of the call's order must exist already
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 16)
    c = torch.zeros((zest_amd.zt_num_coeffs(20), 2), dtype=torch.float64, device=device)
    g = torch.zeros(zest_amd.zt_shape(20), dtype=torch.float64, device=device)
    assert _noalloc(plan, g, c, 20, 1, stream) == 1
    assert b"tables" in lib.zgl_last_error_string()
    # one instance must fit the static array (FP64 Zernike at order 200: ~190 MB)
    big = zest_amd.Plan(zest_amd.ZERNIKE, 200)
    c = torch.zeros((zest_amd.zt_num_coeffs(200), 2), dtype=torch.float64, device=device)
    g = torch.zeros(zest_amd.zt_shape(200), dtype=torch.float64, device=device)
    assert _noalloc(big, g, c, 200, 1, stream) == 1
    assert b"bytes of scratch" in lib.zgl_last_error_string()
    # the same plan still transforms that order through zgl_forward
    big.forward(g, c, 200)
    torch.cuda.synchronize()
