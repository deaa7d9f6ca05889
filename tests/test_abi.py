"""The C ABI boundary without a GPU: the library loads, exports every entry
point include/zgl.h declares, and rejects invalid arguments before touching
the device. Also the host-side helpers (shapes, the portable generator)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "zgl.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zgl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import zest_amd
    names = declared_symbols()
    assert "zgl_forward" in names and "zgl_backward" in names
    for name in names:
        assert hasattr(zest_amd.LIB, name), name
    assert set(zest_amd.EXPORTS) <= set(names)


def test_invalid_arguments_fail_before_device_work():
    import zest_amd
    lib = zest_amd.LIB
    h = ctypes.c_void_p()
    assert lib.zgl_plan_create(None, 0, 1, 0, -1, 0, 4, 0) == 1
    assert lib.zgl_plan_create(ctypes.byref(h), 7, 1, 0, -1, 0, 4, 0) == 1
    assert lib.zgl_plan_create(ctypes.byref(h), 0, 2, 0, -1, 0, 4, 0) == 1
    assert b"zernike_norm" in lib.zgl_last_error_string()
    assert lib.zgl_plan_create(ctypes.byref(h), 0, 1, 5, -1, 0, 4, 0) == 1
    assert lib.zgl_plan_create(ctypes.byref(h), 0, 1, 0, 0, 0, 4, 0) == 1
    assert lib.zgl_plan_create(ctypes.byref(h), 1, 1, 0, -1, 3, 4, 0) == 1
    assert lib.zgl_plan_create(ctypes.byref(h), 0, 1, 0, -1, 0, 4, 9) == 4
    assert lib.zgl_plan_create(ctypes.byref(h), 0, 1, 0, -1, 0, 4, 4) == 4
    assert lib.zgl_plan_create(ctypes.byref(h), 0, 1, 0, -1, 0, 1 << 20, 0) == 1
    assert not h.value
    assert lib.zgl_forward(None, None, 4, None, 4, 1, 0, 0, 0, None) == 1
    assert lib.zgl_backward(None, None, 4, None, 4, 1, 0, 0, 0, None) == 1
    assert lib.zgl_evaluate(None, None, 4, None, 1, None, 1, None, 1, None, 1, 0, 0, 0, None) == 1
    assert lib.zgl_plan_resize(None, 3) == 1
    assert lib.zgl_plan_order(None) == 0
    assert lib.zgl_plan_destroy(None) == 0


def test_shapes_match_zest_layouts():
    import oracle
    import zest_amd
    for order in (0, 1, 2, 7, 8, 32, 128, 255):
        assert zest_amd.zt_shape(order) == oracle.zt_shape(order)
        assert zest_amd.zt_num_coeffs(order) == oracle.zt_num_coeffs(order)
        for layout in (0, 1):
            assert zest_amd.st_shape(order, layout) == oracle.st_shape(order, layout)
    # include/zernike_glq_transformer.hpp:107-126 and include/layout.hpp:400-406
    assert zest_amd.zt_shape(128) == (255, 129, 136)
    assert zest_amd.zt_num_coeffs(32) == 3128 and zest_amd.zt_num_coeffs(128) == 180960
    assert zest_amd.st_shape(512) == (1023, 512) and zest_amd.st_shape(512, 1) == (512, 1024)


def test_synthetic_generator_is_portable():
    from zest_amd import synth
    # SplitMix64 reference outputs for state 0 (Vigna's published generator)
    z = synth.splitmix64(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    u = synth.uniform(5, 10000)
    assert u.min() > 0.0 and u.max() <= 1.0 and abs(u.mean() - 0.5) < 0.01
    x = synth.normal(7, 200000)
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1.0) < 0.01
    c = synth.zernike_coeffs(6, 0)
    n, l, m = synth.zernike_nlm(6)
    assert c.shape == (len(n), 2) and np.all(c[m == 0, 1] == 0.0)
    assert np.array_equal(c, synth.zernike_coeffs(6, 0))
    s = synth.sh_coeffs(6, 1)
    ls, ms = synth.sh_lm(6)
    assert s.shape == (21, 2) and np.all(s[ms == 0, 1] == 0.0)


def test_numpy_arrays_must_be_contiguous_float64():
    import zest_amd
    with pytest.raises(ValueError):
        zest_amd._ptr_and_mem(np.zeros(4, dtype=np.float32))
    with pytest.raises(ValueError):
        zest_amd._ptr_and_mem(np.zeros((4, 4))[:, 0])
    # grids may be float32 (the *_grid32 entry points); coefficients may not
    assert zest_amd._ptr_and_mem(np.zeros(4, dtype=np.float32), allow_f32=True)[2]
    assert not zest_amd._ptr_and_mem(np.zeros(4), allow_f32=True)[2]


def test_euler_angles_from_rotation_matrix_matches_zest():
    """Host helper of the C ABI against zest's constexpr function
    (include/rotor.hpp:49-60), which forms sin(beta) as
    sqrt((1 + r22^2)(1 - r22^2)); exact for beta in {0, pi/2, pi}."""
    import oracle
    import zest_amd
    if oracle.REF is None:
        pytest.skip("zest not compiled (oracle/_ref)")
    rng = np.random.default_rng(1)
    for _ in range(20):
        q, _r = np.linalg.qr(rng.normal(size=(3, 3)))
        q *= np.sign(np.linalg.det(q))
        assert np.abs(zest_amd.euler_angles_from_rotation_matrix(q) - oracle.ref_euler_angles(q)).max() < 1e-15
    a = zest_amd.euler_angles_from_rotation_matrix(np.eye(3))
    assert a[1] == 0.0


def test_binding_rejects_undersized_buffers():
    """The binding checks every buffer against the batch, strides and orders
    it passes to the C ABI and raises before any call (no device needed: an
    order-0 plan builds no tables)."""
    import zest_amd
    zt = zest_amd.Plan(zest_amd.ZERNIKE, 0)
    g, c = np.zeros(zest_amd.zt_shape(8)), np.zeros((zest_amd.zt_num_coeffs(8), 2))
    small_g, small_c = np.zeros(g.size - 1), np.zeros(c.size - 2)
    for call in (lambda: zt.forward(small_g, c, 8), lambda: zt.forward(g, small_c, 8),
                 lambda: zt.backward(small_c, g, 8), lambda: zt.backward(c, small_g, 8),
                 lambda: zt.forward(g, c, 8, batch=2),
                 lambda: zt.backward(c, g, 8, batch=2),
                 lambda: zt.backward_adjoint(g, small_c, 8),
                 lambda: zt.forward_adjoint(small_c, g, 8),
                 lambda: zt.forward(np.zeros(2 * g.size), np.zeros(2 * c.size), 8, batch=2,
                                    grid_stride=g.size + 1),
                 lambda: zt.power_spectrum(small_c, 8),
                 lambda: zt.power_spectrum(c, 8, out=np.zeros(3)),
                 lambda: zt.rotate(small_c, 8, np.zeros(3)),
                 lambda: zt.rotate(c, 8, np.zeros(2)),
                 lambda: zt.polar_rotate(c, 8, np.zeros(1), batch=2),
                 lambda: zt.evaluate(small_c, 8, np.zeros(3), np.zeros(2), np.zeros(2)),
                 lambda: zt.evaluate(c, 8, np.zeros(3), np.zeros(2), np.zeros(2), values=np.zeros(11))):
        with pytest.raises(ValueError):
            call()
    # truncated forward: only the min(exp_order, grid_order) prefix is written
    zt._check_sizes(g, np.zeros((zest_amd.zt_num_coeffs(4), 2)), 8, 4, 1, 0, 0)
    sh = zest_amd.Plan(zest_amd.SH, 0, layout=1)
    gs = np.zeros(zest_amd.st_shape(16, 1))
    with pytest.raises(ValueError):
        sh.forward(gs[:-1], np.zeros((zest_amd.st_num_coeffs(16), 2)), 16)
    with pytest.raises(ValueError):
        sh._jobs([(16, np.zeros(3), np.zeros(3))])


def _no_gpu_visible() -> bool:
    import torch
    return torch.cuda.device_count() == 0


def test_allocator_hook_out_of_memory_on_first_request():
    """An installed device allocator that returns NULL turns the first device
    allocation into ZGL_ERROR_OUT_OF_MEMORY (zest's std::bad_alloc,
    include/alignment.hpp:169-176): no plan is returned, the message names the
    request, nothing is left allocated, and restoring hipMalloc leaves the
    library as it was. No device work is reached: the first allocation of a
    plan comes before its first kernel."""
    import zest_amd
    lib = zest_amd.LIB
    calls = []
    hook = zest_amd.DeviceAllocator(lambda n: calls.append(n) or 0, lambda p, n: None)
    with hook:
        for kind, layout in ((zest_amd.ZERNIKE, 0), (zest_amd.SH, 1)):
            h = ctypes.c_void_p()
            st = lib.zgl_plan_create(ctypes.byref(h), kind, 1, 0, -1, layout, 8, 0)
            assert st == 2, (st, lib.zgl_last_error_string())
            assert not h.value
            msg = lib.zgl_last_error_string().decode()
            assert "installed device allocator returned NULL" in msg and "table allocation" in msg, msg
        # order 0 allocates nothing and succeeds under the failing allocator
        h = ctypes.c_void_p()
        assert lib.zgl_plan_create(ctypes.byref(h), 0, 1, 0, -1, 0, 0, 0) == 0
        assert lib.zgl_plan_destroy(h) == 0
    assert len(calls) == 2 and all(n > 0 for n in calls)
    assert hook.live == 0
    # an allocator without its free function is refused
    assert lib.zgl_set_device_allocator(hook._c_alloc, zest_amd.DEVICE_FREE_FN(), None) == 1
    assert b"free function" in lib.zgl_last_error_string()
    zest_amd.restore_default_allocator()


@pytest.mark.skipif(not _no_gpu_visible(),
                    reason="hands host memory to the library as device memory: only without a GPU")
def test_allocator_hook_releases_everything_when_a_later_request_fails():
    """OOM injected at every point of plan creation in turn: the first k
    requests succeed (host memory, 256-byte aligned; without a GPU no kernel
    runs on it), request k + 1 fails. Whatever the failure point -- and the
    HIP error that ends the sequence once the first table kernel is checked --
    every block that was handed out is returned through the allocator's free,
    with the size it was requested at, and no plan is returned."""
    import zest_amd
    lib = zest_amd.LIB
    seen_statuses = set()
    for kind, order, precision in ((zest_amd.ZERNIKE, 12, 0), (zest_amd.ZERNIKE, 12, 1), (zest_amd.SH, 16, 0)):
        k = 0
        while True:
            held = {}

            def alloc(n, k=k, held=held):
                if len(held) >= k:
                    return 0
                buf = ctypes.create_string_buffer(n + 256)
                addr = (ctypes.addressof(buf) + 255) // 256 * 256
                held[addr] = (buf, n)
                return addr

            def free(p, n, held=held):
                assert p in held and held[p][1] == n
                del held[p]

            hook = zest_amd.DeviceAllocator(alloc, free)
            with hook:
                h = ctypes.c_void_p()
                st = lib.zgl_plan_create(ctypes.byref(h), kind, 1, 0, -1, 0, order, precision)
            assert st in (2, 3), (kind, order, k, st)
            assert not h.value
            assert hook.live == 0 and not held, (kind, order, k, hook.live)
            seen_statuses.add(st)
            if st == 3:  # past the allocations before the first checked kernel
                break
            assert hook.calls == k + 1, (k, hook.calls)
            k += 1
        assert k >= 3, (kind, k)
    assert seen_statuses == {2, 3}
