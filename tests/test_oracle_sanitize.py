"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer:
oracle/sanitize_main.c calls every entry point (FFT sizes 1..1234 in and out
of place, 2D/3D transforms, 1D/2D split runs in both orders and boundaries,
2D/3D Born runs with every preconditioner, cross sections, edge ranges).
Host only."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None,
                    reason="needs gcc and make")
def test_oracle_is_clean_under_asan_and_ubsan(tmp_path):
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize", "SAN_OUT=" + str(tmp_path)],
                   check=True, timeout=300)
    env = dict(os.environ, OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(tmp_path / "orc_sanitize")], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]
    assert "calls returned" in r.stdout
