"""The host code under AddressSanitizer + UndefinedBehaviorSanitizer, on the
CPU: tests/sanitize/host_check (built by tests/sanitize/Makefile from the
library's sources, the C oracle and tests/sanitize/host_check.cpp, with leak
detection on) exercises the C ABI's argument checks, order-0 plans, the Euler
angles, the device-allocator hook with out-of-memory injected at every request
of plan creation, and the oracle's transforms, evaluation and rotations. Any
sanitizer report fails the run. The build's instrumentation is itself checked
with a planted heap overflow in the program (not in any product code)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "sanitize", "_build", "host_check")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _need_exe():
    if not os.path.exists(EXE):
        pytest.skip("sanitizer build absent (make -f tests/sanitize/Makefile)")


def test_host_code_is_clean_under_asan_and_ubsan():
    _need_exe()
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=600, env=ENV)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "0 failures" in r.stdout
    assert "Sanitizer" not in out, out[-4000:]
    print(r.stdout[-1500:])


def test_sanitizer_build_reports_a_planted_overflow():
    _need_exe()
    r = subprocess.run([EXE, "--planted-overflow"], capture_output=True, text=True, timeout=120, env=ENV)
    assert r.returncode != 0
    assert "heap-buffer-overflow" in r.stderr
