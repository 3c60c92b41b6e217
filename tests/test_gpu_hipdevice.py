"""The C++ host adapter: the reference's solver call sequences
(SplittingMethod2D ExpV/ExpT, PreconditionalBornSeries::Update3D) run on
HipDevice / HipFourierTransform*D through the Device plug-in interface, in a
C++ program, compared with the oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "hip_device_driver")


def run(args, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/hip_device_driver not built (run __graft_entry__.build())")
    p = subprocess.run([DRIVER] + [str(a) for a in args], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    return p.stdout


def test_split2d_through_device_interface(tmp_path, oracle):
    nx, ny = 96, 128
    x = -30 + 60.0 / nx * np.arange(nx)
    y = -20 + 40.0 / ny * np.arange(ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    psi = np.exp(0.8j * X - ((X + 5) ** 2 + Y ** 2) / 36.0)
    V = 0.5 * np.exp(-(X ** 2 + Y ** 2) / 8)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(psi.astype(np.complex128).tobytes())
        f.write(V.astype(np.float64).tobytes())
    run(["split2d", nx, ny, -30, 30, -20, 20, 0.05, 30, inp, out], tmp_path)
    got = np.fromfile(out, dtype=np.complex128).reshape(nx, ny)
    ref = oracle.split2d(psi, V, -30.0, 30.0, -20.0, 20.0, 0.05, 30, method=2)
    assert rel_l2(got, ref) < 1e-12


def test_born3d_through_device_interface(tmp_path, oracle):
    n, L, E = 32, 16.0, 3.0
    lo, hi = (-L,) * 3, (L,) * 3
    g = -L + 2 * L / n * np.arange(n)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    V = (0.2 * np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 4)).astype(np.complex128)
    ref, veff, eps, _ = oracle.born(V, lo, hi, E, (0, 0, 1), order=5, preconditional=1,
                                    preconditioner=1, slow=0.5)
    psi0 = oracle.plane_wave((n, n, n), lo, hi, E, (0, 0, 1))
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(psi0.tobytes())
        f.write(veff.tobytes())
    txt = run(["born3d", n, L, E, repr(eps), 0.5, 5, inp, out], tmp_path)
    got = np.fromfile(out, dtype=np.complex128).reshape(n, n, n)
    assert rel_l2(got, ref) < 1e-12
    _, amp = oracle.xsection3d(psi0, ref, veff, lo, hi, E, 0, 0, 1)
    re, im = [float(v) for v in txt.split("xsec_sum")[1].split()]
    assert abs(complex(re, im) - amp) < 1e-11 * abs(amp)
