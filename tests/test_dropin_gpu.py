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
transforms against zest's host-side samplers. They are
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
