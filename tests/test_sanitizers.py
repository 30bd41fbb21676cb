"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the CPU-side code,
each driven over its entry points by a native self-test (no GPU): the
oracle (oracle/or_selftest.c, `make -C oracle sanitize`) and the host-only
parts of the library (tests/native/host_selftest.cpp: the dense-map and PLY
writers, the densify-thread front-end's parallax gate and depth interval,
and the serving scheduler against a simulated executor).  A sanitizer
report or a failed check makes the program exit non-zero."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
           OMP_NUM_THREADS="4")


def _run(cmd, timeout):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=ENV)
    assert r.returncode == 0, f"{cmd}: rc {r.returncode}\n{r.stdout[-3000:]}\n{r.stderr[-6000:]}"
    return r.stdout


def test_oracle_under_asan_ubsan():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], 300)
    out = _run([os.path.join(ROOT, "oracle", "build", "or_selftest_asan")], 300)
    assert "or_selftest: ok" in out


def test_host_code_under_asan_ubsan():
    native = os.path.join(ROOT, "tests", "native")
    if not os.path.exists(os.path.join(native, "Makefile")):
        pytest.skip("no host self-test")
    _run(["make", "-s", "-C", native], 300)
    out = _run([os.path.join(native, "build", "host_selftest_asan")], 300)
    assert "host_selftest: ok" in out
