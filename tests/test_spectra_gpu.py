"""zgl_power_spectrum / zgl_cross_power_spectrum (zest's power_spectra.hpp) on
the GPU against the oracle, which is pinned against compiled zest and
Parseval's identity. The per-entry sums run in zest's order; tolerance 1e-14
relative to the largest entry (FMA contraction may differ)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("kind", ["zernike", "sh"])
@pytest.mark.parametrize("order", [1, 2, 17, 128])
@pytest.mark.parametrize("sn", [0, 1])
def test_power_and_cross_match_oracle(device, kind, order, sn):
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    zt = kind == "zernike"
    gen = synth.zernike_coeffs if zt else synth.sh_coeffs
    ref_fn = oracle.zt_power_spectrum if zt else oracle.st_power_spectrum
    batch = 3
    a = np.stack([gen(order, 10 + i) for i in range(batch)])
    b = np.stack([gen(order, 20 + i) for i in range(batch)])
    plan = zest_amd.Plan(zest_amd.ZERNIKE if zt else zest_amd.SH, 4, sh_norm=sn)
    da, db = torch.from_numpy(a).to(device), torch.from_numpy(b).to(device)
    p = plan.power_spectrum(da, order, batch=batch).cpu().numpy()
    x = plan.power_spectrum(da, order, batch=batch, other=db).cpu().numpy()
    for i in range(batch):
        assert rel(p[i], ref_fn(a[i], order, sn)) < 1e-14
        assert rel(x[i], ref_fn(a[i], order, sn, b[i])) < 1e-14
    # host memory gives the same numbers
    assert np.array_equal(plan.power_spectrum(a, order, batch=batch), p)
    assert np.array_equal(plan.power_spectrum(a, order, batch=batch, other=b), x)


def test_strides(device):
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    order, batch = 9, 4
    nc = 2 * zest_amd.zt_num_coeffs(order)
    nent = (order + 1) ** 2 // 4
    a = np.zeros((batch, nc + 6))
    for i in range(batch):
        a[i, :nc] = synth.zernike_coeffs(order, i).ravel()
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 4)
    out = torch.full((batch, nent + 3), -5.0, dtype=torch.float64, device=device)
    plan.power_spectrum(torch.from_numpy(a).to(device), order, out=out, batch=batch,
                        coeff_stride=nc + 6, out_stride=nent + 3)
    o = out.cpu().numpy()
    assert (o[:, nent:] == -5.0).all()
    for i in range(batch):
        assert rel(o[i, :nent], oracle.zt_power_spectrum(a[i, :nc], order)) < 1e-14


@pytest.mark.parametrize("shift,pad", [(1, 0), (0, 3), (1, 5)])
def test_unaligned_instances(device, shift, pad):
    """Instance bases off 16-byte alignment (a base shifted by one double, or an
    odd stride in doubles) take the 8-byte-load form; aligned ones the 16-byte
    one. Both match the oracle, power and cross spectra."""
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    order, batch = 33, 3
    nc = 2 * zest_amd.zt_num_coeffs(order)
    stride = nc + pad
    nent = (order + 1) ** 2 // 4
    bufs = []
    for seed in (0, 50):
        flat = np.zeros(shift + batch * stride)
        for i in range(batch):
            flat[shift + i * stride: shift + i * stride + nc] = synth.zernike_coeffs(order, seed + i).ravel()
        bufs.append(flat)
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 4)
    da = torch.from_numpy(bufs[0]).to(device)[shift:]
    db = torch.from_numpy(bufs[1]).to(device)[shift:]
    p = plan.power_spectrum(da, order, batch=batch, coeff_stride=stride).cpu().numpy()
    x = plan.power_spectrum(da, order, batch=batch, coeff_stride=stride, other=db).cpu().numpy()
    for i in range(batch):
        a_i = bufs[0][shift + i * stride: shift + i * stride + nc]
        b_i = bufs[1][shift + i * stride: shift + i * stride + nc]
        assert rel(p[i, :nent], oracle.zt_power_spectrum(a_i, order)) < 1e-14
        assert rel(x[i, :nent], oracle.zt_power_spectrum(a_i, order, 0, b_i)) < 1e-14
