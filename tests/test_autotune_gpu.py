"""The autotuner (zgl_plan_autotune): the first call at an order times the
candidate FFT tile widths and instances per workgroup and keeps the fastest.
The choice must not change any result: each instance's arithmetic is the
same whatever the grouping, so tuned and untuned plans agree bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_tuned_results_identical(device, precision):
    import torch
    import zest_amd
    from zest_amd import synth
    out = {}
    for tuned in (False, True):
        plan = zest_amd.Plan(zest_amd.ZERNIKE, 8, precision=precision)
        plan.autotune(tuned)
        res = []
        for order, batch in ((24, 6), (64, 5)):
            c = np.stack([synth.zernike_coeffs(order, 90 + b) for b in range(batch)])
            dc = torch.from_numpy(c).to(device)
            dg = torch.empty((batch,) + zest_amd.zt_shape(order), dtype=torch.float64, device=device)
            co = torch.zeros_like(dc)
            for _ in range(2):  # the second call runs the tuned choices
                plan.backward(dc, dg, order, batch=batch)
                plan.forward(dg, co, order, batch=batch)
            torch.cuda.synchronize()
            res.append((dg.cpu().numpy(), co.cpu().numpy()))
        out[tuned] = (res, plan.tuning())
    for (g0, c0), (g1, c1) in zip(out[False][0], out[True][0]):
        assert np.array_equal(g0, g1) and np.array_equal(c0, c1)
    assert out[False][1] == []
    tuning = out[True][1]
    print("tuning:", tuning)
    stages = {(o, st) for o, _, _, st, _ in tuning}
    for order in (24, 64):
        for st in ("fft_forward", "legendre_forward", "radial_forward", "radial_backward",
                   "legendre_backward", "fft_backward"):
            assert (order, st) in stages
    for o, rb, gb, st, choice in tuning:
        assert rb == (8 if precision == 0 else 4) and gb == 8
        assert choice in ((16, 8, 4, 2, 1) if st.startswith("fft") else (1, 2, 4, 8))
