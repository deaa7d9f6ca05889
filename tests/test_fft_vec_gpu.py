"""The 16-byte-per-lane memory phases of the longitudinal FFT kernels (used
for Zernike grids whose instance bases are 16-byte aligned) against the
one-value-per-lane kernels (ZGL_FFT_SCALAR=1): the arithmetic is the same,
so the results are bit-identical, in every precision, over FFT tile widths,
prime and composite lengths, batches and strides, and ragged jobs. A
misaligned batch (an odd instance stride) takes the one-value-per-lane form
and still agrees."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(device, order, batch, precision, grid32, stride_pad, seed):
    import torch
    import zest_amd
    from zest_amd import synth
    c = np.stack([synth.zernike_coeffs(order, seed + s) for s in range(batch)])
    dc = torch.from_numpy(c).to(device)
    shape = zest_amd.zt_shape(order)
    gsz = int(np.prod(shape))
    gdt = torch.float32 if grid32 else torch.float64
    st = gsz + stride_pad
    buf = torch.zeros(batch * st, dtype=gdt, device=device)
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order, precision=precision)
    plan.backward(dc, buf, order, batch=batch, grid_stride=st)
    co = torch.zeros_like(dc)
    plan.forward(buf, co, order, batch=batch, grid_stride=st)
    torch.cuda.synchronize()
    flat = buf.cpu().numpy()
    grids = np.stack([flat[b * st:b * st + gsz].reshape(shape) for b in range(batch)])
    return grids, co.cpu().numpy()


@pytest.mark.parametrize("order,batch", [(7, 3), (16, 5), (33, 4), (64, 3), (96, 2), (128, 3)])
@pytest.mark.parametrize("precision,grid32", [(0, False), (1, False), (1, True), (2, False)])
def test_vector_memory_phases_bit_identical(device, monkeypatch, order, batch, precision, grid32):
    monkeypatch.delenv("ZGL_FFT_SCALAR", raising=False)
    g_vec, c_vec = _pair(device, order, batch, precision, grid32, 0, 5)
    monkeypatch.setenv("ZGL_FFT_SCALAR", "1")
    g_sc, c_sc = _pair(device, order, batch, precision, grid32, 0, 5)
    assert np.array_equal(g_vec, g_sc)
    assert np.array_equal(c_vec, c_sc)


@pytest.mark.parametrize("tile", ["8", "16", "32"])
def test_vector_phases_each_tile_width(device, monkeypatch, tile):
    monkeypatch.setenv("ZGL_FFT_TILE", tile)
    for precision, grid32 in ((0, False), (1, True)):
        monkeypatch.delenv("ZGL_FFT_SCALAR", raising=False)
        a = _pair(device, 40, 3, precision, grid32, 0, 9)
        monkeypatch.setenv("ZGL_FFT_SCALAR", "1")
        b = _pair(device, 40, 3, precision, grid32, 0, 9)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_misaligned_stride_takes_scalar_path(device, monkeypatch):
    """An instance stride of an odd number of values leaves the second
    instance 8 (FP64) or 4 (FP32) bytes off 16-byte alignment."""
    import oracle
    monkeypatch.delenv("ZGL_FFT_SCALAR", raising=False)
    for precision, grid32 in ((0, False), (1, True)):
        g, c = _pair(device, 24, 3, precision, grid32, 1, 21)
        monkeypatch.setenv("ZGL_FFT_SCALAR", "1")
        g2, c2 = _pair(device, 24, 3, precision, grid32, 1, 21)
        monkeypatch.delenv("ZGL_FFT_SCALAR", raising=False)
        assert np.array_equal(c, c2)
        ref = oracle.zt_backward(c[1], 24)
        tol = 1e-12 if precision == 0 else 1e-6
        assert np.abs(g[1] - ref).max() <= tol * np.abs(ref).max()

