"""Host check (CPU) of the LDS placements of the FFT line exchanges
(qsim_fft_core.h Pow2Line::pad): for the padding, the block swizzles and the
per-line key, every element of a line must land in a slot of its own inside
the line's LS_RAW slots, for every radix a plan exchanges after and every
key.  A swizzle that is not a permutation of its block, or a placement that
outgrows the line's slots, would corrupt the transforms only on the GPU;
here it fails on the host."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_every_placement_is_a_permutation_within_the_line(tmp_path):
    exe = str(tmp_path / "placement_check")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O1", "-std=c++17",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "qusim_amd", "csrc"),
           "-x", "hip", os.path.join(ROOT, "tests", "cpp", "placement_check.cpp"), "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "placements ok" in r.stdout
