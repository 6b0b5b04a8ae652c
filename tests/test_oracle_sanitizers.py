"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer:
oracle/san_check.c runs every oracle function on small lattices (all four
stencils, both dimensions, the string finder at several thresholds with
full and cut lists, sums, transforms and spectra) in a build with
-fsanitize=address,undefined; any out-of-bounds access, leak or undefined
arithmetic ends the run with the sanitizer's report."""
import os
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def test_oracle_runs_clean_under_asan_and_ubsan():
    b = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], capture_output=True, text=True,
                       timeout=600)
    if b.returncode != 0 and "asan" in (b.stderr + b.stdout).lower():
        pytest.skip("the sanitizer runtime is not available: " + b.stderr[-300:])
    assert b.returncode == 0, b.stdout + b.stderr
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([os.path.join(ORACLE, "_san", "san_check")], capture_output=True,
                       text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr
    assert "oracle sanitizer run: ok" in r.stdout
