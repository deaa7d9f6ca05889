"""zest's own transformer and grid-evaluator test programs, compiled unchanged
against the drop-in headers (include/zest_amd/) and libzgl.so, run on the GPU.

The programs are zest's tests/test_zernike_glq_transformer.cpp (known-answer
forward and backward transforms of single basis functions at order 6 and an
order-100 round trip at 1e-13, all for (normed, geo, none) and
(unnormed, qm, cs)) and tests/test_sh_glq_transformer.cpp (the same for
spherical harmonics over both grid layouts and all four conventions) and
tests/test_grid_evaluator.cpp (constant, Y_10, Y_31 + Y_4,-3, Z_33,-2 + Z_531
on grids through both poles, r = 0 and r = 1, at 1e-13, for every
convention combination) and tests/test_rotor.cpp (the d(pi/2) matrices to
l = 4 against closed forms, rotations of SH, Zernike and row-skipping
Zernike-SH spans, identity rotations at 1e-13). device_sampling_test (a HIP
program of this repository, tests/dropin/device_sampling_test.hip) samples
functions on the quadrature grids on the device with
include/zest_amd/device_sampling.hpp and checks the grids and their forward
transforms against zest's host-side samplers. row_skipping_backward_test
(tests/dropin/row_skipping_backward_test.cpp) calls the row-skipping SH
backward overload on Zernike rows and writes the grids, which the test below
compares with zest's own overload; rotate_l_test (tests/dropin/rotate_l_test.cpp)
does the same for the per-degree rotations rotate_l / polar_rotate_l. They are
built by tests/dropin/Makefile where the zest sources are present; the test
skips when the binaries were not built.
"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dropin", "_bin")


@pytest.mark.parametrize("name", ["zest_test_zernike_glq_transformer", "zest_test_sh_glq_transformer",
                                  "zest_test_grid_evaluator", "zest_test_rotor",
                                  "device_sampling_test"])
def test_zest_transformer_tests_pass_on_gpu(name):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip("drop-in test programs not built (zest sources absent at build time)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, f"{name} failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"


def test_row_skipping_sh_backward_through_dropin(tmp_path):
    """zest's row-skipping SH backward overload (include/sh_glq_transformer.hpp:
    680-700), called on rows of a Zernike expansion through the drop-in header
    by tests/dropin/row_skipping_backward_test.cpp (all four SH conventions,
    both grid layouts, grid orders equal to, above and below the row's order,
    parity-changing truncations included). Each grid is compared with zest's
    own overload (oracle/_ref) where it was compiled, else with the oracle's
    backward transform of the expansion that overload reads, at 1e-12
    relative to the grid's largest value."""
    import numpy as np
    import oracle
    exe = os.path.join(BIN, "row_skipping_backward_test")
    if not os.path.exists(exe):
        pytest.skip("drop-in test programs not built (zest sources absent at build time)")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, f"failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
    files = sorted(tmp_path.glob("case_*.bin"))
    assert len(files) == 96, r.stdout
    zest_t = {}
    worst = 0.0
    for path in files:
        raw = path.read_bytes()
        head = np.frombuffer(raw[:64], dtype=np.int64)
        zorder, n, order, sn, sp, layout, ncoef, ngrid = (int(x) for x in head)
        body = np.frombuffer(raw[64:], dtype=np.float64)
        assert body.size == ncoef + ngrid
        c = body[:ncoef].reshape(-1, 2)
        g = body[ncoef:].reshape(oracle.st_shape(order, layout))
        if oracle.REF is not None:
            key = (sn, sp, layout)
            if key not in zest_t:
                zest_t[key] = oracle.ZestSH(20, sn, sp, layout)
            ref = zest_t[key].backward_zernike_row(c, zorder, n, order)
        else:
            e = oracle.st_row_skipping_expansion(c, zorder, n, order)
            ref = oracle.st_backward(e, order, min(n + 1, order), sh_norm=sn, sh_phase=sp, layout=layout)
        err = float(np.abs(g - ref).max() / np.abs(ref).max())
        worst = max(worst, err)
        assert err < 1e-12, (path.name, n, order, sn, sp, layout, err)
    print(f"{len(files)} row-skipping backward cases, max rel err {worst:.2e}")


def test_rotate_l_through_dropin(tmp_path):
    """zest::rotate_l / zest::polar_rotate_l (include/wignerd_collection.hpp:
    122, 304) from the drop-in header, run by tests/dropin/rotate_l_test.cpp:
    the program checks whole SH (both convention pairs) and Zernike rotations
    composed of them, as zest's Rotor composes them, against the drop-in Rotor
    on the GPU at 1e-12, and zero angles against the identity; here each
    per-degree result (l = 0 .. 127, random vectors and phase factors) is
    compared with zest's own functions at 1e-14 relative to the largest value."""
    import numpy as np
    import oracle
    exe = os.path.join(BIN, "rotate_l_test")
    if not os.path.exists(exe):
        pytest.skip("drop-in test programs not built (zest sources absent at build time)")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, f"failed ({r.returncode}):\n{r.stdout[-3000:]}\n{r.stderr[-2000:]}"
    print(r.stdout)
    files = sorted(tmp_path.glob("rl_*.bin"))
    assert len(files) == 10
    if oracle.REF is None:
        pytest.skip("zest not compiled (oracle/_ref): per-degree comparison not run")
    for path in files:
        raw = path.read_bytes()
        l = int(np.frombuffer(raw[:8], dtype=np.int64)[0])
        v, eg, eb, ea, rot, pol = np.frombuffer(raw[16:], dtype=np.complex128).reshape(6, l + 1)
        ref = oracle.ref_rotate_l(v, eg, eb, ea)
        assert np.abs(rot - ref).max() <= 1e-14 * np.abs(ref).max(), l
        ref = oracle.ref_polar_rotate_l(v, ea)
        assert np.abs(pol - ref).max() <= 1e-14 * np.abs(ref).max(), l
