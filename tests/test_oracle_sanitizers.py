"""The CPU oracle under AddressSanitizer + UndefinedBehaviorSanitizer and ThreadSanitizer.

oracle/sanitize.mk links oracle/sanitize_check.c (checks of every oracle entry point on
small inputs that reach its edge paths: identity rotations, deflation, block limits, odd
orders, n = 0 and 1, padded leading dimensions, the iteration limit) with the oracle
source, once per sanitizer; each binary exits non-zero on a sanitizer report or a failed
check.  The oracle is the parity checker of the HIP library, so a memory error or
undefined behaviour in it would void the parity tests; the HIP library itself is not
built with sanitizers (GPU code cannot be).
"""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
BINARIES = ["check_asan", "check_asan_fast", "check_tsan"]


@pytest.fixture(scope="module")
def sanitized():
    if shutil.which(os.environ.get("CC", "gcc")) is None or shutil.which("make") is None:
        pytest.skip("no C compiler or make")
    r = subprocess.run(["make", "-s", "-j3", "-f", "sanitize.mk"], cwd=ORACLE,
                       capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        if "asan" in r.stderr.lower() or "tsan" in r.stderr.lower():
            pytest.skip("sanitizer runtime not available: " + r.stderr[-300:])
        raise AssertionError("sanitizer build failed:\n" + r.stderr[-2000:])
    return os.path.join(ORACLE, "_san")


@pytest.mark.parametrize("name", BINARIES)
def test_oracle_sanitizer_clean(sanitized, name):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["TSAN_OPTIONS"] = "halt_on_error=1"
    r = subprocess.run([os.path.join(sanitized, name)], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, f"{name} exited {r.returncode}:\n{r.stderr[-4000:]}"
    assert "0 Float64 and 0 Float32 failures" in r.stdout
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr
