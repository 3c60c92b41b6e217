"""The evolver's and the scattering solver's observables, the FNES step and
the complex (absorbing) potential on the GPU, against the golden vectors of
the compiled reference (tests/golden/ref_vectors.npz) and the FP64 oracle.

Bar: complex128 psi within 1e-12 relative L2; scalar observables within
1e-12 relative (reductions), cross sections within 1e-10 relative."""
import math
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.npz")


@pytest.fixture(scope="module")
def G():
    return np.load(GOLD, allow_pickle=False)


def evolver1d(psi, V, dt, batch=1, method=2, boundary="Period", fnes=False, complex_v=False,
              mass=1.0, hbar=1.0, precision="fp64", x0=-100.0, x1=100.0):
    from qusim_amd.api import BoundaryCondition, Evolver1D, Options, SolverMethod
    o = Options().Hip().Batch(batch)
    if complex_v:
        o.ComplexPotential(True)
    if precision != "fp64":
        o.Precision(precision)
    return Evolver1D().init(psi, False, dt, fnes, V, x0, x1, psi.size,
                            getattr(BoundaryCondition, boundary),
                            SolverMethod.SplittingMethodO2 if method == 2
                            else SolverMethod.SplittingMethodO4, mass, hbar, o)


def evolver2d(psi, V, dt, batch=1, method=2, fnes=False, complex_v=False, mass=1.0, hbar=1.0,
              ext=(-30.0, 30.0, -20.0, 25.0)):
    from qusim_amd.api import BoundaryCondition, Evolver2D, Options, SolverMethod
    o = Options().Hip().Batch(batch)
    if complex_v:
        o.ComplexPotential(True)
    return Evolver2D().init(psi, False, dt, fnes, V, ext[0], ext[1], psi.shape[0], ext[2],
                            ext[3], psi.shape[1], BoundaryCondition.Period,
                            SolverMethod.SplittingMethodO2 if method == 2
                            else SolverMethod.SplittingMethodO4, mass, hbar, o)


# ------------------------------------------------------- complex potential
def test_complex_potential_1d_matches_reference_vectors(G, oracle):
    """The C1 grid with -i 0.1 (d/20)^2 absorbing edges through the fused 1D
    kernel (2048 points), O2 for 1000 steps and O4 for 150, against the
    compiled reference's complex MulExp trajectories."""
    psi, V = G["c1_psi0"], G["c1cv_V"]
    e = evolver1d(psi, V, 0.5, batch=50, complex_v=True)
    e.steps(20)
    assert rel_l2(e.GetPsi(), G["c1cv_o2_1000"]) < 1e-12
    e4 = evolver1d(psi, V, 0.5, batch=50, method=4, complex_v=True)
    e4.steps(3)
    assert rel_l2(e4.GetPsi(), G["c1cv_o4_150"]) < 1e-12
    # FP32-mixed on the same path: complex64 psi, FP64 factors
    e32 = evolver1d(psi, V, 0.5, batch=50, complex_v=True, precision="mixed32")
    e32.steps(3)
    err = rel_l2(e32.GetPsi(), G["c1cv_o2_150"])
    print("FP32-mixed 1D absorbing, 150 steps: rel L2", err)
    assert err < 1e-5


def test_complex_potential_2d_matches_reference_vectors(G, oracle):
    """64 x 48 with an absorbing layer (the unfused reference order, since 48
    is not a power of two), O2 and O4; and 64 x 64 (fused sweeps) against the
    oracle, whose complex branch the same vectors pin."""
    p2, V2 = G["s2_psi0"], G["s2cv_V"]
    e = evolver2d(p2, V2, 0.1, batch=80, complex_v=True)
    e.step()
    assert rel_l2(e.GetPsi(), G["s2cv_o2_80"]) < 1e-12
    e4 = evolver2d(p2, V2, 0.1, batch=30, method=4, complex_v=True)
    e4.step()
    assert rel_l2(e4.GetPsi(), G["s2cv_o4_30"]) < 1e-12
    n = 64
    x = -30 + 60.0 / n * np.arange(n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    p = np.exp(0.8j * X - ((X + 5) ** 2 + Y ** 2) / 36.0)
    d = np.minimum.reduce([X + 30.0, 30.0 - X, Y + 30.0, 30.0 - Y])
    V = 0.5 * np.exp(-(X ** 2 + Y ** 2) / 8) - 1j * np.where(d < 8.0, 0.2 * (d / 8.0) ** 2, 0.0)
    for m in (2, 4):
        ef = evolver2d(p, V, 0.1, batch=60, method=m, complex_v=True,
                       ext=(-30.0, 30.0, -30.0, 30.0))
        ef.step()
        ref = oracle.split2d(p, V, -30.0, 30.0, -30.0, 30.0, 0.1, 60, method=m, complex_v=True)
        assert rel_l2(ef.GetPsi(), ref) < 1e-12


def test_absorbing_norm_non_increasing_on_device(G):
    psi, V = G["c1_psi0"], G["c1cv_V"]
    e = evolver1d(psi, V, 0.5, batch=10, complex_v=True)
    norms = [e.Norm2()]
    for _ in range(30):
        e.step()
        norms.append(e.Norm2())
    assert np.all(np.diff(norms) <= 1e-14)
    assert norms[-1] < 0.5 * norms[0]


# ------------------------------------------------------------------- FNES
def test_fnes_matches_reference_vectors(G):
    """Step() renormalises after its `batch` split steps (src/EvolverImpl.cpp:
    56-68): 1D fused (2048) and 2D unfused (64 x 48), both with absorption."""
    e = evolver1d(G["c1_psi0"], G["c1cv_V"], 0.5, batch=2, fnes=True, complex_v=True)
    for _ in range(120):
        e.step()
    assert rel_l2(e.GetPsi(), G["c1cv_fnes_o2_120x2"]) < 1e-12
    assert abs(e.Norm2() - 1.0) < 1e-13
    e2 = evolver2d(G["s2_psi0"], G["s2cv_V"], 0.1, batch=2, fnes=True, complex_v=True)
    e2.steps(40)
    assert rel_l2(e2.GetPsi(), G["s2cv_fnes_o2_40x2"]) < 1e-12


# ------------------------------------------------------------ observables
def test_observables_match_reference_values(G, oracle):
    """KinEn, PotEn, Norm2, Xavg of the evolver (device FFT plans and
    compensated device reductions) against the reference's device-op
    sequences, for the periodic and the infinite-wall 1D evolver and the 2D
    one; also against the oracle."""
    x = -100 + 200.0 / 2048 * np.arange(2048)
    V = 0.6 * np.exp(-x ** 2 / 2)
    cases = (("obs_c1_periodic", G["c1cv_o2_150"], "Period", 1.0, 1.0),
             ("obs_c1_wall", G["c1_wall_o2_200"], "InfiniteWall", 1.0, 1.0),
             ("obs_c1_hbar_mass", G["c1cv_o2_150"], "InfiniteWall", 1.7, 0.8))
    for key, s, b, m, hb in cases:
        e = evolver1d(s, V, 0.5, boundary=b, mass=m, hbar=hb)
        got = (e.KinEn(), e.PotEn(), e.Norm2(), e.Xavg())
        want = G[key]
        for g, w, what in zip(got, want, ("KinEn", "PotEn", "Norm2", "Xavg")):
            assert g == pytest.approx(w, rel=1e-12, abs=1e-14), (key, what)
        o = oracle.observables(s, V, -100.0, 100.0, mass=m, hbar=hb,
                               boundary=0 if b == "Period" else 1)
        assert np.allclose(got, o, rtol=1e-12, atol=1e-14), key
    s2, V2 = G["s2cv_o2_80"], G["s2_V"]
    e2 = evolver2d(s2, V2, 0.1, mass=1.3, hbar=0.9)
    got2 = (e2.KinEn(), e2.PotEn(), e2.Norm2())
    for g, w in zip(got2, G["obs_s2"][:3]):
        assert g == pytest.approx(w, rel=1e-12)


def test_free_packet_observables_closed_forms():
    """Gaussian packet: <T> = (k^2 + 1/(4 s^2)) / 2m, <x> = c, PotEn of x^2
    = c^2 + s^2; and <x> moves at velocity k/m under free evolution
    (applications/QuEvol1DPrecision/QuSimPrecision.cpp:210-218)."""
    n, L, k0, c, s, m = 4096, 80.0, 1.3, -7.0, 2.5, 1.7
    x = -L + 2 * L / n * np.arange(n)
    psi = np.exp(1j * k0 * x - (x - c) ** 2 / (4 * s * s))
    e = evolver1d(psi, x ** 2, 0.01, batch=100, mass=m, x0=-L, x1=L)
    assert e.KinEn() == pytest.approx((k0 ** 2 + 1 / (4 * s * s)) / (2 * m), rel=1e-12)
    assert e.PotEn() == pytest.approx(c * c + s * s, rel=1e-12)
    assert e.Xavg() == pytest.approx(c, rel=1e-12)
    ef = evolver1d(psi, np.zeros(n), 0.02, batch=250, mass=m, x0=-L, x1=L)
    x_start = ef.Xavg()
    ef.step()
    assert (ef.Xavg() - x_start) / ef.Time() == pytest.approx(k0 / m, rel=1e-9)
    assert ef.KinEn() == pytest.approx((k0 ** 2 + 1 / (4 * s * s)) / (2 * m), rel=1e-12)


# -------------------------------------------------------- cross sections
def test_total_cross_sections_match_oracle(oracle):
    """ComputeTotalXSection 2D (src/ScatteringSolverImpl.cpp:199-212) and 3D
    (:363-379, theta by j) over the device cross-section reduction."""
    from qusim_amd.api import (Options, QuPerturbation2D, QuPerturbation3D, SolverMethod)
    o = Options().Hip().Order(6).Preconditional(True).VellekoopPreconditioner().Slow(0.7)
    lo, hi, sh = (-60.0, -50.0), (60.0, 50.0), (96, 80)
    xs = [lo[i] + (hi[i] - lo[i]) / sh[i] * np.arange(sh[i]) for i in range(2)]
    X, Y = np.meshgrid(*xs, indexing="ij")
    V = 0.3 * np.exp(-(X ** 2 + Y ** 2) / 4) + 0.05j * np.exp(-(X ** 2 + Y ** 2) / 9)
    s = QuPerturbation2D().init(V, lo[0], hi[0], sh[0], lo[1], hi[1], sh[1], 2.0, 0.01, 1.0, 0.3,
                                SolverMethod.BornSerise, 1.0, 1.0, o)
    s.Compute()
    ref, veff, _, _ = oracle.born(V, lo, hi, 2.0, (1.0, 0.3), order=6, preconditional=1,
                                  preconditioner=1, slow=0.7, epsilon=0.01)
    psi0 = oracle.plane_wave(sh, lo, hi, 2.0, (1.0, 0.3))
    want = oracle.total_xsection2d(psi0, ref, veff, lo, hi, 2.0, 24)
    assert s.ComputeTotalXSection(24) == pytest.approx(want, rel=1e-10)
    sh3 = (32, 24, 28)
    lo3 = tuple(-q / 2.0 for q in sh3)
    hi3 = tuple(q / 2.0 for q in sh3)
    g3 = [lo3[i] + (hi3[i] - lo3[i]) / sh3[i] * np.arange(sh3[i]) for i in range(3)]
    X3, Y3, Z3 = np.meshgrid(*g3, indexing="ij")
    V3 = (0.2 * np.exp(-(X3 ** 2 + Y3 ** 2 + Z3 ** 2) / 4)).astype(np.complex128)
    o3 = Options().Hip().Order(5).Preconditional(True).VellekoopPreconditioner().Slow(0.5)
    s3 = QuPerturbation3D().init(V3, lo3[0], hi3[0], sh3[0], lo3[1], hi3[1], sh3[1], lo3[2],
                                 hi3[2], sh3[2], 4.0, 0.01, 0.0, 0.2, 1.0,
                                 SolverMethod.BornSerise, 1.0, 1.0, o3)
    s3.Compute()
    ref3, veff3, _, _ = oracle.born(V3, lo3, hi3, 4.0, (0.0, 0.2, 1.0), order=5,
                                    preconditional=1, preconditioner=1, slow=0.5, epsilon=0.01,
                                    threads=8)
    psi03 = oracle.plane_wave(sh3, lo3, hi3, 4.0, (0.0, 0.2, 1.0))
    want3 = oracle.total_xsection3d(psi03, ref3, veff3, lo3, hi3, 4.0, 6, 10)
    assert s3.ComputeTotalXSection(6, 10) == pytest.approx(want3, rel=1e-10)


def test_delta_norm_2d_on_device_matches_oracle(oracle):
    """GetDeltaPsiNorm after each Compute (src/Perturbation2DImpl.cpp:80-82),
    reduced on the device, for three successive calls."""
    from qusim_amd.api import Options, QuPerturbation2D, SolverMethod
    lo, hi, sh = (-60.0, -50.0), (60.0, 50.0), (128, 128)
    xs = [lo[i] + (hi[i] - lo[i]) / sh[i] * np.arange(sh[i]) for i in range(2)]
    X, Y = np.meshgrid(*xs, indexing="ij")
    V = 0.3 * np.exp(-(X ** 2 + Y ** 2) / 4) + 0.05j * np.exp(-(X ** 2 + Y ** 2) / 9)
    o = Options().Hip().Order(4).Preconditional(True).Hao1Preconditioner().Slow(0.7)
    s = QuPerturbation2D().init(V, lo[0], hi[0], sh[0], lo[1], hi[1], sh[1], 2.0, 0.01, 1.0, 0.3,
                                SolverMethod.BornSerise, 1.0, 1.0, o)
    for calls in (1, 2, 3):
        s.Compute()
        _, _, _, dn = oracle.born(V, lo, hi, 2.0, (1.0, 0.3), order=4, ncalls=calls,
                                  preconditional=1, preconditioner=2, slow=0.7, epsilon=0.01)
        assert s.GetDeltaPsiNorm() == pytest.approx(dn, rel=1e-10), calls
    assert math.isfinite(s.GetDeltaPsiNorm())


# ------------------------------------------------------------ imaginary time
def test_imaginary_time_ground_state_on_device(oracle):
    """Complex dt = -0.05 i with FNES after every call of 10 steps: the fused
    1D kernel (256 points) and the fused 2D sweeps (128 x 96) relax a moving,
    displaced packet to the oscillator ground state; psi matches the oracle
    run with the same calls to 1e-12, and the device KinEn + PotEn is the
    ground-state energy w / 2 ((wx + wy) / 2) to 1e-6."""
    n, L, w = 256, 20.0, 1.3
    x = -L / 2 + L / n * np.arange(n)
    V = 0.5 * w * w * x * x
    psi = np.exp(-(x - 1.5) ** 2 / 2 + 0.7j * x)
    e = evolver1d(psi, V, -0.05j, batch=10, fnes=True, x0=-L / 2, x1=L / 2)
    for _ in range(30):
        e.step()
    p0 = psi / np.sqrt(np.sum(np.abs(psi) ** 2) * L / n)
    ref = oracle.split1d(p0, V, -L / 2, L / 2, -0.05j, 30, method=2, batch=10, fnes=True)
    assert rel_l2(e.GetPsi(), ref) < 1e-12
    assert e.KinEn() + e.PotEn() == pytest.approx(w / 2, abs=1e-6)
    nx, ny, Lx, Ly, wx, wy = 128, 96, 16.0, 14.0, 1.0, 1.6
    gx = -Lx / 2 + Lx / nx * np.arange(nx)
    gy = -Ly / 2 + Ly / ny * np.arange(ny)
    X, Y = np.meshgrid(gx, gy, indexing="ij")
    V2 = 0.5 * (wx * wx * X * X + wy * wy * Y * Y)
    p2 = np.exp(-((X - 1) ** 2 + (Y + 0.5) ** 2) / 3 + 0.4j * X)
    e2 = evolver2d(p2, V2, -0.05j, batch=10, fnes=True, ext=(-Lx / 2, Lx / 2, -Ly / 2, Ly / 2))
    e2.steps(30)
    q0 = p2 / np.sqrt(np.sum(np.abs(p2) ** 2) * (Lx / nx) * (Ly / ny))
    ref2 = oracle.split2d(q0, V2, -Lx / 2, Lx / 2, -Ly / 2, Ly / 2, -0.05j, 30, method=2,
                          batch=10, fnes=True)
    assert rel_l2(e2.GetPsi(), ref2) < 1e-12
    assert e2.KinEn() + e2.PotEn() == pytest.approx((wx + wy) / 2, abs=1e-6)
