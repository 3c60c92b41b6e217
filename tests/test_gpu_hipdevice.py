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


OP_NAMES = ["SetZero", "SetOne", "Exp(complex)", "Exp(real)", "ExpI", "Copy",
            "Scale(double) in place", "Scale(complex) in place", "Scale(double) out",
            "Scale(complex) out", "Sub", "Mul", "MulR", "MulMinusI", "Mul(real, complex)",
            "Mul(real, alpha, complex)", "Mul in place", "MulExp out", "MulExp(complex)",
            "MulExp(real)", "MulExpI", "CopyReverseMinu", "MulExpIdx2", "MulExpK1D",
            "MulExpK2D", "MulExpK3D", "G01D", "G02D", "G03D", "AddMul", "Add3 in place",
            "Add3 out", "LinearUpdate", "Vellekoop", "Hao1", "Hao2", "Copy(real)",
            "Scale(real) in place", "Scale(real) out", "Set(complex)", "Set(real)"]
SCALAR_NAMES = ["Abs2Idx(complex)", "Abs2Idx(real)", "At(complex)", "At(real)", "Min", "Max",
                "SumReal", "SumImag", "MinusNorm2", "Norm2(real)", "Norm2(complex)", "Abs2Idx2",
                "Abs2K1D", "Abs2K2D", "Abs2K3D", "Abs2Mul(complex)", "Abs2Mul(real)",
                "Xsection3D"]


def test_every_device_op_matches_the_reference(tmp_path):
    """Every elementwise op and reduction of the Device interface, run by
    HipDevice on the battery's fixed inputs (oracle/device_op_battery.h),
    against the same battery on the compiled reference's CpuSeq device
    (golden vectors): elementwise ops to 1e-14 relative L2, reductions to
    1e-12.  MulExpK3D is the stated divergence: the reference drops alpha
    (src/DeviceTemplate.h:876), HipDevice applies it."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "ref_vectors.npz"))
    dims = [int(d) for d in g["ops_dims"]]
    n = int(np.prod(dims))
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        for k in ("ops_a", "ops_b", "ops_c"):
            f.write(g[k].astype(np.complex128).tobytes())
        f.write(g["ops_r"].astype(np.float64).tobytes())
    run(["ops"] + dims + [inp, out], tmp_path)
    got = np.fromfile(out, dtype=np.complex128)
    arrays, scalars = got[:len(OP_NAMES) * n].reshape(len(OP_NAMES), n), got[len(OP_NAMES) * n:]
    want_a, want_s = g["ops_arrays"], g["ops_scalars"]
    assert want_a.shape == (len(OP_NAMES), n) and want_s.shape == (len(SCALAR_NAMES),)
    for i, name in enumerate(OP_NAMES):
        if name == "MulExpK3D":
            continue
        if not np.any(want_a[i]):
            assert not np.any(arrays[i]), name
        else:
            assert rel_l2(arrays[i], want_a[i]) < 1e-14, (name, rel_l2(arrays[i], want_a[i]))
    for i, name in enumerate(SCALAR_NAMES):
        assert abs(scalars[i] - want_s[i]) <= 1e-12 * abs(want_s[i]), (name, scalars[i], want_s[i])
    # MulExpK3D: psi * exp(alpha (i^2 ax + j^2 ay + k^2 az)) with folded indices
    nx, ny, nz = dims
    fold = lambda i, m: np.where(i > m // 2, m - i, i)  # noqa: E731
    I, J, K = np.meshgrid(fold(np.arange(nx), nx), fold(np.arange(ny), ny),
                          fold(np.arange(nz), nz), indexing="ij")
    expo = (0.2 + 0.1j) * (I ** 2 * 0.03 + J ** 2 * 0.05 + K ** 2 * 0.07)
    k3 = g["ops_a"] * np.exp(expo.ravel())
    assert rel_l2(arrays[OP_NAMES.index("MulExpK3D")], k3) < 1e-14
    assert rel_l2(want_a[OP_NAMES.index("MulExpK3D")], k3) > 1e-3
