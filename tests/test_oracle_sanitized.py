"""The oracle under AddressSanitizer and UBSan: `make -C oracle selftest`
builds oracle/selftest.c (which includes oracle.c) with
-fsanitize=address,undefined, and the self-test runs 300 random products,
transposes, shrink maps and binary COO round trips against dense brute
force; any out-of-bounds access, leak or undefined behaviour fails it."""
import os
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def test_oracle_selftest_sanitized(tmp_path):
    build = subprocess.run(["make", "-s", "-C", ORACLE, "selftest"], capture_output=True, text=True)
    if build.returncode != 0:
        pytest.skip("sanitizer build unavailable: " + build.stderr[-300:])
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([os.path.join(ORACLE, "_build", "selftest"), str(tmp_path / "selftest.coo")],
                         capture_output=True, text=True, env=env, timeout=300)
    assert run.returncode == 0, run.stderr[-3000:]
    assert "300 trials passed" in run.stdout
