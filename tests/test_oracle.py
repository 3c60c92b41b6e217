"""CPU pins of the oracle (oracle/qsim_oracle.c).  Each check is independent of
the oracle's own formulas: golden vectors of the compiled reference, closed
forms, invariants, brute force, and the reference's published known answers.
A planted mistake in the oracle (a sign, an index, a dropped term) fails one."""
import math
import os

import numpy as np
import pytest

from conftest import rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "ref_vectors.npz")


@pytest.fixture(scope="module")
def G():
    return np.load(GOLD, allow_pickle=False)


def ramp(shape):
    return np.arange(int(np.prod(shape)), dtype=np.float64).reshape(shape).astype(np.complex128)


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 12, 16, 17, 64, 100, 243, 1234])
@pytest.mark.parametrize("inverse", [False, True])
def test_fft_vs_bruteforce(oracle, n, inverse):
    x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
    assert rel_l2(oracle.fft(x, inverse), oracle.dft_bruteforce(x, inverse)) < 2e-15


def test_fft_closed_forms(oracle):
    n = 96
    d = np.zeros(n, np.complex128)
    d[0] = 1
    assert np.allclose(oracle.fft(d), np.ones(n), atol=1e-15)          # delta -> constant
    k0 = 5
    e = np.exp(2j * np.pi * k0 * np.arange(n) / n)                      # exponential -> delta
    want = np.zeros(n, np.complex128)
    want[k0] = n
    assert np.max(np.abs(oracle.fft(e) - want)) < 1e-12
    x = np.random.default_rng(0).standard_normal((12, 23, 34)).astype(np.complex128)
    assert rel_l2(oracle.fft(oracle.fft(x), True), x * x.size) < 1e-14  # round trip
    y = oracle.fft(x)
    assert abs(np.sum(np.abs(y) ** 2) / x.size - np.sum(np.abs(x) ** 2)) < 1e-10  # Parseval


@pytest.mark.parametrize("shape", [(123, 234), (12, 23, 34), (64, 48)])
def test_fft_multi_d_vs_numpy(oracle, shape):
    x = np.random.default_rng(3).standard_normal(shape) + 0j
    assert rel_l2(oracle.fft(x), np.fft.fftn(x)) < 1e-14
    assert rel_l2(oracle.fft(x, True), np.fft.ifftn(x) * x.size) < 1e-14


def test_fft_golden(oracle, G):
    for name, shape in (("fft1234", (1234,)), ("fft123x234", (123, 234)),
                        ("fft12x23x34", (12, 23, 34)), ("fft64x48", (64, 48))):
        x = ramp(shape)
        f, i = oracle.fft(x), oracle.fft(x, True)
        if name + "_idx" in G:
            idx = G[name + "_idx"]
            assert rel_l2(f.ravel()[idx], G[name + "_fwd_sample"]) < 1e-14, name
            assert rel_l2(i.ravel()[idx], G[name + "_inv_sample"]) < 1e-14, name
            s = G[name + "_sums"]
            assert abs(np.sum(np.abs(f) ** 2) - s[1].real) < 1e-13 * s[1].real, name
        else:
            assert rel_l2(f, G[name + "_fwd"]) < 1e-14, name
            assert rel_l2(i, G[name + "_inv"]) < 1e-14, name


# -------------------------------------------------------------- split step
def test_split_golden(oracle, G):
    psi = G["c1_psi0"]
    x = -100 + 200.0 / 2048 * np.arange(2048)
    V = 0.6 * np.exp(-x ** 2 / 2)
    assert rel_l2(oracle.split1d(psi, V, -100.0, 100.0, 0.5, 1000, method=2), G["c1_o2_1000"]) < 1e-12
    assert rel_l2(oracle.split1d(psi, V, -100.0, 100.0, 0.5, 200, method=4), G["c1_o4_200"]) < 1e-12
    assert rel_l2(oracle.split1d(psi, V, -100.0, 100.0, 0.5, 200, method=2, boundary=1),
                  G["c1_wall_o2_200"]) < 1e-12
    assert rel_l2(oracle.split1d(psi, V, -100.0, 100.0, 0.5 - 0.05j, 100, method=2),
                  G["c1_cdt_o2_100"]) < 1e-12
    p2, V2 = G["s2_psi0"], G["s2_V"]
    assert rel_l2(oracle.split2d(p2, V2, -30.0, 30.0, -20.0, 25.0, 0.1, 50, method=2),
                  G["s2_o2_50"]) < 1e-12
    assert rel_l2(oracle.split2d(p2, V2, -30.0, 30.0, -20.0, 25.0, 0.1, 20, method=4),
                  G["s2_o4_20"]) < 1e-12


def _packet(n, L=50.0, k=1.3, c=-10.0):
    x = -L + 2 * L / n * np.arange(n)
    return x, np.exp(1j * k * x - (x - c) ** 2 / 20)


@pytest.mark.parametrize("n", [512, 243])
def test_free_evolution_closed_form(oracle, n):
    """V = 0: the split step is exactly exp(-i k^2 T/2) in k space (odd n pins
    the frequency folding of the Nyquist neighbourhood)."""
    T, steps = 3.0, 60
    x, psi = _packet(n)
    k = 2 * np.pi * np.fft.fftfreq(n, d=100.0 / n)
    want = np.fft.ifft(np.exp(-0.5j * k * k * T) * np.fft.fft(psi))
    for m in (2, 4):
        got = oracle.split1d(psi, np.zeros(n), -50.0, 50.0, T / steps, steps, method=m)
        assert rel_l2(got, want) < 1e-12


@pytest.mark.parametrize("shape", [(128, 96), (75, 121)])
def test_free_evolution_closed_form_2d(oracle, shape):
    T, steps = 3.0, 60
    L = (-40.0, 40.0)
    xs = [L[0] + (L[1] - L[0]) / s * np.arange(s) for s in shape]
    X, Y = np.meshgrid(*xs, indexing="ij")
    p2 = np.exp(1j * X - (X ** 2 + Y ** 2) / 20)
    kx = 2 * np.pi * np.fft.fftfreq(shape[0], d=(L[1] - L[0]) / shape[0])
    ky = 2 * np.pi * np.fft.fftfreq(shape[1], d=(L[1] - L[0]) / shape[1])
    KX, KY = np.meshgrid(kx, ky, indexing="ij")
    want2 = np.fft.ifft2(np.exp(-0.5j * (KX ** 2 + KY ** 2) * T) * np.fft.fft2(p2))
    for m in (2, 4):
        got2 = oracle.split2d(p2, np.zeros(shape), L[0], L[1], L[0], L[1], T / steps, steps,
                              method=m)
        assert rel_l2(got2, want2) < 1e-12


def test_norm_conservation_and_time_reversal(oracle):
    n = 1024
    x, psi = _packet(n)
    V = 0.5 * np.exp(-x ** 2)
    fwd = oracle.split1d(psi, V, -50.0, 50.0, 0.05, 400, method=2)
    assert abs(oracle.norm2(fwd) / oracle.norm2(psi) - 1) < 1e-13
    back = oracle.split1d(fwd, V, -50.0, 50.0, -0.05, 400, method=2)
    assert rel_l2(back, psi) < 1e-12


def test_convergence_orders(oracle):
    """error(dt) ~ dt^2 for O2 and ~ dt^4 for O4 against a fine reference"""
    n, T = 512, 2.0
    x, psi = _packet(n, k=1.0, c=-3.0)
    V = 1.0 * np.exp(-x ** 2)
    ref = oracle.split1d(psi, V, -50.0, 50.0, T / 2048, 2048, method=4)
    for m, order in ((2, 2), (4, 4)):
        e1 = rel_l2(oracle.split1d(psi, V, -50.0, 50.0, T / 32, 32, method=m), ref)
        e2 = rel_l2(oracle.split1d(psi, V, -50.0, 50.0, T / 64, 64, method=m), ref)
        assert abs(math.log2(e1 / e2) - order) < 0.3, (m, e1, e2)


def test_tunnelling_known_answer(oracle):
    """applications/QuEvol1DPrecision/QuSimPrecision.cpp:82: P(x > 0) at T = 64
    for gauss(x,-20,5) e^{ix}, V = exp(-x^2) on [-100,100) = 0.15937744085."""
    n, x0, x1, dt = 2000, -100.0, 100.0, 0.04
    dx = (x1 - x0) / n
    x = dx * np.arange(n) + x0
    g = np.exp(-(x + 20) ** 2 / 50.0) / (5 * math.sqrt(2 * math.pi))  # Cal gauss = normal pdf
    psi = g * np.exp(1j * x)
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * dx)
    p = oracle.split1d(psi, np.exp(-x * x), x0, x1, dt, 1600, method=4)
    assert abs(oracle.norm2_range_1d(p, x0, x1, 0.0, 100.0) - 0.15937744085) < 5e-8


def test_free_packet_velocity(oracle):
    """applications/QuEvol1DPrecision/QuSimPrecision.cpp:210-218: (Xavg(T)-Xavg(0))/T = 1"""
    n, x0, x1 = 4000, -100.0, 100.0
    dx = (x1 - x0) / n
    x = dx * np.arange(n) + x0
    psi = np.exp(-x ** 2 / 50.0) * np.exp(1j * x)
    T, steps = 20.0, 400
    p = oracle.split1d(psi, np.zeros(n), x0, x1, T / steps, steps, method=2)

    def xavg(q):
        w = np.abs(q) ** 2
        return x0 + dx * np.sum(w * np.arange(n)) / np.sum(w)

    assert abs((xavg(p) - xavg(psi)) / T - 1) < 1e-6


# ------------------------------------------------------------------ Born
def test_born_golden(oracle, G):
    lo2, hi2 = (-60.0, -50.0), (60.0, 50.0)
    for tag, pc, pre, slow in (("naive", 0, 0, 1.0), ("ident", 1, 0, 0.6), ("vell", 1, 1, 0.7),
                               ("hao1", 1, 2, 0.7), ("hao2", 1, 3, 0.5)):
        dp, veff, eps, dn = oracle.born(G["b2_V"], lo2, hi2, 2.0, (1.0, 0.3), order=3, ncalls=2,
                                        preconditional=pc, preconditioner=pre, slow=slow,
                                        epsilon=0.01)
        assert rel_l2(dp, G["b2_%s_dpsi" % tag]) < 1e-13, tag
        eg, dg = G["b2_%s_scalars" % tag]
        assert abs(eps - eg) <= 1e-13 * abs(eg) and abs(dn - dg) <= 1e-11 * abs(dg), tag
        if tag == "vell":
            assert rel_l2(veff, G["b2_vell_veff"]) == 0.0
    lo3, hi3 = (-6.0, -5.0, -7.0), (6.0, 5.0, 7.0)
    for tag, pc, pre, slow in (("naive", 0, 0, 1.0), ("vell", 1, 1, 0.5), ("hao2", 1, 3, 0.5)):
        dp, veff, eps, dn = oracle.born(G["b3_V"], lo3, hi3, 4.0, (0.0, 0.2, 1.0), order=3,
                                        preconditional=pc, preconditioner=pre, slow=slow,
                                        epsilon=0.01)
        assert rel_l2(dp, G["b3_%s_dpsi" % tag]) < 1e-13, tag
        if tag == "vell":
            psi0 = oracle.plane_wave(G["b3_V"].shape, lo3, hi3, 4.0, (0.0, 0.2, 1.0))
            xs, _ = oracle.xsection3d(psi0, dp, veff, lo3, hi3, 4.0, 0.3, 0.1, 0.9)
            assert abs(xs - G["b3_vell_xsec"][0]) < 1e-12 * xs


def test_first_born_closed_form_3d(oracle):
    """applications/QuScattering3DPrecision/QuScattering3DPrecision.cpp:11-21:
    V = v0 exp(-a r^2) gives dsigma/dOmega = m^2/(4 pi^2 hbar^4) [v0 (pi/a)^1.5 e^{-q^2/4a}]^2."""
    n, L, v0, a, p = 48, 8.0, 0.01, 1.0, 1.0
    g = -L + 2 * L / n * np.arange(n)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    V = (v0 * np.exp(-a * (X ** 2 + Y ** 2 + Z ** 2))).astype(np.complex128)
    E = p * p / 2
    psi0 = oracle.plane_wave(V.shape, (-L,) * 3, (L,) * 3, E, (0, 0, 1))
    for th in (0.0, 0.7, 2.0):
        c = (math.sin(th) * math.cos(0.1234), math.sin(th) * math.sin(0.1234), math.cos(th))
        got, _ = oracle.xsection3d(psi0, np.zeros_like(V), V, (-L,) * 3, (L,) * 3, E, *c)
        q2 = (c[0] ** 2 + c[1] ** 2 + (c[2] - 1) ** 2) * p * p
        mat = v0 * (math.pi / a) ** 1.5 * math.exp(-q2 / (4 * a))
        want = mat * mat / (2 * math.pi * 2 * math.pi)
        assert abs(got - want) < 1e-10 * want


def test_first_born_closed_form_2d(oracle):
    """applications/QuScattering2DPrecision/QuScattering2DPrecision.cpp:5-14."""
    n, L, v0, a, p = 256, 20.0, 0.001, 1.0, 1.0
    g = -L + 2 * L / n * np.arange(n)
    X, Y = np.meshgrid(g, g, indexing="ij")
    V = (v0 * np.exp(-a * (X ** 2 + Y ** 2))).astype(np.complex128)
    E = 0.5
    psi0 = oracle.plane_wave(V.shape, (-L, -L), (L, L), E, (1, 0))
    for th in (0.3, 1.5, 3.0):
        got = oracle.xsection2d(psi0, np.zeros_like(V), V, (-L, -L), (L, L), E, math.cos(th),
                                math.sin(th))
        kx, ky = (1 - math.cos(th)) * p, math.sin(th) * p
        mat = v0 * math.pi * math.exp(-(kx * kx + ky * ky) / (4 * a)) / a
        want = mat * mat / (2 * math.pi * p)
        assert abs(got - want) < 1e-10 * want


def test_born_fixed_point_is_lippmann_schwinger(oracle):
    """The converged naive series solves (E - T_k + i eps) dpsi_k = FFT(V (psi0 + dpsi))."""
    n, L, E, eps = 24, 12.0, 2.0, 0.05
    g = -L + 2 * L / n * np.arange(n)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    V = (0.05 * np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 3)).astype(np.complex128)
    dp, _, _, _ = oracle.born(V, (-L,) * 3, (L,) * 3, E, (0, 0, 1), order=80,
                              preconditional=0, epsilon=eps)
    psi0 = oracle.plane_wave(V.shape, (-L,) * 3, (L,) * 3, E, (0, 0, 1))
    k = 2 * np.pi * np.fft.fftfreq(n, d=2 * L / n)
    KX, KY, KZ = np.meshgrid(k, k, k, indexing="ij")
    lhs = (E - (KX ** 2 + KY ** 2 + KZ ** 2) / 2 + 1j * eps) * np.fft.fftn(dp)
    rhs = np.fft.fftn(V * (psi0 + dp))
    assert rel_l2(lhs, rhs) < 1e-10


def test_epsilon_rule(oracle):
    """GetEpsilon (src/Perburbation.h:329-344): eps = max(eps_user, max|V_eff|) for
    Vellekoop; 0.5 |V|^2 / (-sum Im V) for Identity."""
    n = 16
    lo, hi = (-8.0,) * 3, (8.0,) * 3
    V = np.full((n, n, n), 0.1 + 0.0j)
    _, veff, eps, _ = oracle.born(V, lo, hi, 4.0, (0, 0, 1), order=1, preconditional=1,
                                  preconditioner=1, epsilon=0.0)
    assert eps == pytest.approx(np.max(np.abs(veff)), rel=1e-15)
    _, veff0, eps0, _ = oracle.born(V, lo, hi, 4.0, (0, 0, 1), order=1, preconditional=1,
                                    preconditioner=0, epsilon=0.0)
    assert eps0 == pytest.approx(0.5 * np.sum(np.abs(veff0) ** 2) / -np.sum(veff0.imag), rel=1e-13)


def test_compiled_reference_agrees_when_present(oracle):
    """With the reference tree present, the oracle equals the compiled reference
    on fresh random inputs too (not only on the committed vectors)."""
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(11)
    x = rng.standard_normal((20, 30)) + 1j * rng.standard_normal((20, 30))
    assert rel_l2(oracle.fft(x), oracle.ref_fft(x)) < 1e-14
    n = 300
    xs, psi = _packet(n)
    V = rng.uniform(0, 1, n) * 0.3
    assert rel_l2(oracle.split1d(psi, V, -50.0, 50.0, 0.07, 120, method=4),
                  oracle.ref_split1d(psi, V, -50.0, 50.0, 0.07, 120, method=4)) < 1e-12
    # the committed Device-op battery is what the compiled reference gives
    g = np.load(GOLD)
    arrays, scalars = oracle.ref_device_ops(tuple(int(d) for d in g["ops_dims"]), g["ops_a"],
                                            g["ops_b"], g["ops_c"], g["ops_r"])
    assert np.array_equal(arrays, g["ops_arrays"]) and np.array_equal(scalars, g["ops_scalars"])


# ------------------------------------------------ complex potential (C2 / C3)
def test_complex_potential_golden(oracle, G):
    """The imaginary absorbing layer of the 2D workloads runs the oracle's
    complex branch of exp(-i dt t V / hbar).  Golden trajectories come from
    the reference's own complex MulExp / Exp device ops (src/Device.h:104,129)
    and Step()'s FNES renormalisation (src/EvolverImpl.cpp:60-63)."""
    psi, V = G["c1_psi0"], G["c1cv_V"]
    run = lambda steps, **kw: oracle.split1d(psi, V, -100.0, 100.0, 0.5, steps,  # noqa: E731
                                             complex_v=True, **kw)
    assert rel_l2(run(1000), G["c1cv_o2_1000"]) < 1e-12
    assert rel_l2(run(150), G["c1cv_o2_150"]) < 1e-12
    assert rel_l2(run(150, method=4), G["c1cv_o4_150"]) < 1e-12
    assert rel_l2(run(120, batch=2, fnes=True), G["c1cv_fnes_o2_120x2"]) < 1e-12
    p2, V2 = G["s2_psi0"], G["s2cv_V"]
    run2 = lambda steps, **kw: oracle.split2d(p2, V2, -30.0, 30.0, -20.0, 25.0, 0.1,  # noqa: E731
                                              steps, complex_v=True, **kw)
    assert rel_l2(run2(80), G["s2cv_o2_80"]) < 1e-12
    assert rel_l2(run2(30, method=4), G["s2cv_o4_30"]) < 1e-12
    assert rel_l2(run2(40, batch=2, fnes=True), G["s2cv_fnes_o2_40x2"]) < 1e-12


@pytest.mark.parametrize("method", [2, 4])
def test_constant_complex_potential_closed_form(oracle, method):
    """V = V0 - i W constant: the potential factor commutes with the kinetic
    one, so psi(T) = exp(-i V0 T) exp(-W T) x free evolution, exactly.  Pins
    both parts of alpha * V (a sign slip in either shows as a phase or a
    growth instead of a decay)."""
    n, T, steps, V0, W = 256, 4.0, 40, 0.3, 0.05
    x, psi = _packet(n)
    k = 2 * np.pi * np.fft.fftfreq(n, d=100.0 / n)
    free = np.fft.ifft(np.exp(-0.5j * k * k * T) * np.fft.fft(psi))
    want = np.exp(-1j * V0 * T) * math.exp(-W * T) * free
    V = np.full(n, V0 - 1j * W)
    got = oracle.split1d(psi, V, -50.0, 50.0, T / steps, steps, method=method, complex_v=True)
    assert rel_l2(got, want) < 1e-12
    shape = (48, 40)
    xs = [-20.0 + 40.0 / s * np.arange(s) for s in shape]
    X, Y = np.meshgrid(*xs, indexing="ij")
    p2 = np.exp(1j * X - (X ** 2 + Y ** 2) / 20)
    kx = 2 * np.pi * np.fft.fftfreq(shape[0], d=40.0 / shape[0])
    ky = 2 * np.pi * np.fft.fftfreq(shape[1], d=40.0 / shape[1])
    KX, KY = np.meshgrid(kx, ky, indexing="ij")
    want2 = (np.exp(-1j * V0 * T) * math.exp(-W * T)
             * np.fft.ifft2(np.exp(-0.5j * (KX ** 2 + KY ** 2) * T) * np.fft.fft2(p2)))
    got2 = oracle.split2d(p2, np.full(shape, V0 - 1j * W), -20.0, 20.0, -20.0, 20.0, T / steps,
                          steps, method=method, complex_v=True)
    assert rel_l2(got2, want2) < 1e-12


def test_absorbing_layer_norm_is_non_increasing(oracle, G):
    """Im V <= 0 makes every potential factor a contraction and the kinetic
    factor is unitary, so the norm never grows from one step to the next; it
    falls strictly once the packet reaches the layer (2D: the C2 construction
    on a small grid)."""
    psi, V = G["c1_psi0"], G["c1cv_V"]
    norms = [oracle.norm2(psi)]
    p = psi
    for _ in range(60):
        p = oracle.split1d(p, V, -100.0, 100.0, 0.5, 5, complex_v=True)
        norms.append(oracle.norm2(p))
    d = np.diff(norms)
    assert np.all(d <= 1e-14 * norms[0])
    assert norms[-1] < 0.5 * norms[0]  # the packet (v = 1, from x = -30) crossed x = 80
    p2, V2 = G["s2_psi0"], G["s2cv_V"]
    norms2 = [oracle.norm2(p2)]
    q = p2
    for _ in range(20):
        q = oracle.split2d(q, V2, -30.0, 30.0, -20.0, 25.0, 0.1, 10, complex_v=True)
        norms2.append(oracle.norm2(q))
    assert np.all(np.diff(norms2) <= 1e-14 * norms2[0])
    assert norms2[-1] < norms2[0]


def test_fnes_equals_normalised_plain_run(oracle, G):
    """The step is linear in psi, so renormalising after every Step() gives
    the plain run divided by its own norm (FNES, src/EvolverImpl.cpp:60-63)."""
    psi, V = G["c1_psi0"], G["c1cv_V"]
    dx = 200.0 / 2048
    plain = oracle.split1d(psi, V, -100.0, 100.0, 0.5, 240, complex_v=True)
    fnes = oracle.split1d(psi, V, -100.0, 100.0, 0.5, 120, batch=2, complex_v=True, fnes=True)
    assert rel_l2(fnes, plain / math.sqrt(oracle.norm2(plain) * dx)) < 1e-12
    assert abs(oracle.norm2(fnes) * dx - 1) < 1e-13
    p2, V2 = G["s2_psi0"], G["s2cv_V"]
    dv = (60.0 / 64) * (45.0 / 48)
    plain2 = oracle.split2d(p2, V2, -30.0, 30.0, -20.0, 25.0, 0.1, 80, complex_v=True)
    fnes2 = oracle.split2d(p2, V2, -30.0, 30.0, -20.0, 25.0, 0.1, 40, batch=2, complex_v=True,
                           fnes=True)
    assert rel_l2(fnes2, plain2 / math.sqrt(oracle.norm2(plain2) * dv)) < 1e-12


# ---------------------------------------------------- evolver observables
def test_observables_golden(oracle, G):
    """KinEn / PotEn / Norm2 / Xavg against the reference's device-op
    sequences (Abs2K1D, Abs2K2D, Abs2Idx2, Abs2Mul, Abs2Idx, Norm2)."""
    x = -100 + 200.0 / 2048 * np.arange(2048)
    V = 0.6 * np.exp(-x ** 2 / 2)
    cases = (("obs_c1_periodic", G["c1cv_o2_150"], dict(boundary=0)),
             ("obs_c1_wall", G["c1_wall_o2_200"], dict(boundary=1)),
             ("obs_c1_hbar_mass", G["c1cv_o2_150"], dict(boundary=1, mass=1.7, hbar=0.8)))
    for key, s, kw in cases:
        got = oracle.observables(s, V, -100.0, 100.0, **kw)
        assert np.allclose(got, G[key], rtol=1e-12, atol=1e-13), key
    got = oracle.observables(G["s2cv_o2_80"], G["s2_V"], -30.0, 30.0, -20.0, 25.0, mass=1.3,
                             hbar=0.9)
    assert np.allclose(got[:3], G["obs_s2"][:3], rtol=1e-12, atol=1e-13)


def test_observables_closed_forms(oracle):
    """A Gaussian packet exp(i k x - (x-c)^2 / (4 s^2)) on a fine periodic grid:
    <k^2> = k^2 + 1/(4 s^2), <x> = c, <x^2> = c^2 + s^2 (spectrally exact).
    The 1D KinEn of the reference divides by 2m without hbar
    (src/SplittingMethod1D.cpp:103-104), the 2D one uses (hbar Dk)^2 / 2m."""
    n, L, k0, c, s, m = 4096, 80.0, 1.3, -7.0, 2.5, 1.7
    x = -L + 2 * L / n * np.arange(n)
    psi = np.exp(1j * k0 * x - (x - c) ** 2 / (4 * s * s))
    kin, pot, nrm, xavg = oracle.observables(psi, x ** 2, -L, L, mass=m, hbar=0.6)
    assert kin == pytest.approx((k0 ** 2 + 1 / (4 * s * s)) / (2 * m), rel=1e-12)
    assert pot == pytest.approx(c * c + s * s, rel=1e-12)
    assert xavg == pytest.approx(c, rel=1e-12)
    assert nrm == pytest.approx(math.sqrt(2 * math.pi) * s, rel=1e-12)
    # infinite wall: sin(j pi q / N) is an eigenvector of the wall's sine
    # transform, KinEn = (hbar q pi / L)^2 / 2m exactly; a two-mode mixture
    # gives the |amplitude|^2-weighted mean
    N, q1, q2, hb = 1000, 3, 7, 0.8
    j = np.arange(N)
    a, b = 1.0, 0.5j
    w = a * np.sin(np.pi * q1 * j / N) + b * np.sin(np.pi * q2 * j / N)
    Lw = 20.0
    e = lambda q: (hb * q * math.pi / Lw) ** 2 / (2 * m)  # noqa: E731
    want = (abs(a) ** 2 * e(q1) + abs(b) ** 2 * e(q2)) / (abs(a) ** 2 + abs(b) ** 2)
    kin_w = oracle.observables(w, np.zeros(N), 0.0, Lw, mass=m, hbar=hb, boundary=1)[0]
    assert kin_w == pytest.approx(want, rel=1e-12)
    # 2D: separable packet, kinetic energy adds over the axes
    nx, ny, sx, sy, kx, ky, hb2 = 256, 192, 2.0, 3.0, 0.7, -0.4, 0.9
    gx = -30 + 60.0 / nx * np.arange(nx)
    gy = -36 + 72.0 / ny * np.arange(ny)
    X, Y = np.meshgrid(gx, gy, indexing="ij")
    p2 = np.exp(1j * (kx * X + ky * Y) - X ** 2 / (4 * sx * sx) - Y ** 2 / (4 * sy * sy))
    kin2, pot2, _, _ = oracle.observables(p2, X ** 2 + 2 * Y ** 2, -30.0, 30.0, -36.0, 36.0,
                                          mass=m, hbar=hb2)
    want2 = hb2 ** 2 * (kx * kx + ky * ky + 1 / (4 * sx * sx) + 1 / (4 * sy * sy)) / (2 * m)
    assert kin2 == pytest.approx(want2, rel=1e-12)
    assert pot2 == pytest.approx(sx * sx + 2 * sy * sy, rel=1e-12)


# ------------------------------------------------------ total cross sections
def _first_born_3d(v0, a, p, cth):
    q2 = 2 * p * p * (1 - cth)
    mat = v0 * (math.pi / a) ** 1.5 * math.exp(-q2 / (4 * a))
    return mat * mat / (4 * math.pi * math.pi)


def test_total_xsection_closed_forms(oracle):
    """ComputeTotalXSection (src/ScatteringSolverImpl.cpp:199-212, 363-379) of a
    weak Gaussian: (1) equal to the same angular sums of the first-Born closed
    form dsigma/dOmega; (2) close to the closed-form integral, since both
    angular rules are accurate for this smooth integrand (the 2D rule is
    periodic; the 3D theta rule has zero end weights through sin(theta))."""
    from scipy.special import i0e
    # 3D, theta indexed by j (the reference's i at :369 is a documented slip)
    n, L, v0, a, p = 40, 8.0, 0.01, 1.0, 1.0
    g = -L + 2 * L / n * np.arange(n)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    V = (v0 * np.exp(-a * (X ** 2 + Y ** 2 + Z ** 2))).astype(np.complex128)
    E = p * p / 2
    psi0 = oracle.plane_wave(V.shape, (-L,) * 3, (L,) * 3, E, (0, 0, 1))
    npsi, nth = 6, 24
    got = oracle.total_xsection3d(psi0, np.zeros_like(V), V, (-L,) * 3, (L,) * 3, E, npsi, nth)
    disc = sum(_first_born_3d(v0, a, p, math.cos(j * math.pi / nth)) * math.sin(j * math.pi / nth)
               for j in range(nth) for i in range(npsi)) * 4 * math.pi / (nth * npsi)
    assert got == pytest.approx(disc, rel=1e-9)
    pref = (v0 * (math.pi / a) ** 1.5 / (2 * math.pi)) ** 2
    integral = pref * 2 * math.pi * (a / (p * p)) * (1 - math.exp(-2 * p * p / a))
    # The reference weights each sample by 4 pi / (ntheta npsi) (:377), which
    # sums sin(theta_j) over the sphere to 8, not 4 pi: the rectangle rule's
    # weight is 2 pi^2 / (ntheta npsi).  Parity keeps the reference's weight,
    # so the total is (2 / pi) x the integral.
    assert got == pytest.approx(integral * 2 / math.pi, rel=5e-3)
    # 2D
    n2, L2, v2 = 192, 16.0, 0.001
    g2 = -L2 + 2 * L2 / n2 * np.arange(n2)
    X2, Y2 = np.meshgrid(g2, g2, indexing="ij")
    V2 = (v2 * np.exp(-a * (X2 ** 2 + Y2 ** 2))).astype(np.complex128)
    psi02 = oracle.plane_wave(V2.shape, (-L2, -L2), (L2, L2), E, (1, 0))
    ns = 16
    got2 = oracle.total_xsection2d(psi02, np.zeros_like(V2), V2, (-L2, -L2), (L2, L2), E, ns)
    c2 = (v2 * math.pi / a) ** 2 / (2 * math.pi * p)
    disc2 = sum(c2 * math.exp(-p * p * (1 - math.cos(2 * math.pi * i / ns)) / a)
                for i in range(ns)) * 2 * math.pi / ns
    assert got2 == pytest.approx(disc2, rel=1e-9)
    # int_0^{2 pi} exp(-p^2 (1 - cos t) / a) dt = 2 pi e^{-p^2/a} I0(p^2/a)
    assert got2 == pytest.approx(c2 * 2 * math.pi * i0e(p * p / a), rel=1e-9)


def test_imaginary_time_reaches_the_oscillator_ground_state(oracle):
    """Complex dt = -i tau with FNES (src/EvolverImpl.cpp:56-68) is
    imaginary-time propagation: a displaced, moving packet in
    V = (w x)^2 / 2 (m = hbar = 1) relaxes to the ground state
    exp(-w x^2 / 2) with energy w / 2 (2D: (wx + wy) / 2).  Strang splitting
    in imaginary time has no negative sub-steps; the fourth-order scheme's
    negative ones (D2 < 0, src/SplittingUtils.cpp:3-13) amplify high
    momenta, so only O2 is a ground-state solver."""
    n, L, w = 256, 20.0, 1.3
    x = -L / 2 + L / n * np.arange(n)
    dx = L / n
    V = 0.5 * w * w * x * x
    psi = np.exp(-(x - 1.5) ** 2 / 2 + 0.7j * x)
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * dx)
    g = np.exp(-w * x * x / 2)
    g /= np.sqrt(np.sum(g * g) * dx)
    out = oracle.split1d(psi, V, -L / 2, L / 2, -0.05j, 300, method=2, fnes=True)
    kin, pot, nrm, xavg = oracle.observables(out, V, -L / 2, L / 2)
    assert nrm == pytest.approx(1.0, abs=1e-13)
    assert kin + pot == pytest.approx(w / 2, abs=1e-6)
    # virial theorem <T> = <V>; the splitting's fixed point shifts each by O(tau^2)
    assert kin == pytest.approx(w / 4, abs=1e-3) and pot == pytest.approx(w / 4, abs=1e-3)
    assert abs(np.sum(np.conj(g) * out) * dx) ** 2 > 1 - 1e-6
    assert abs(xavg) < 1e-6
    nx, ny, Lx, Ly, wx, wy = 128, 96, 16.0, 14.0, 1.0, 1.6
    gx = -Lx / 2 + Lx / nx * np.arange(nx)
    gy = -Ly / 2 + Ly / ny * np.arange(ny)
    X, Y = np.meshgrid(gx, gy, indexing="ij")
    V2 = 0.5 * (wx * wx * X * X + wy * wy * Y * Y)
    p2 = np.exp(-((X - 1) ** 2 + (Y + 0.5) ** 2) / 3 + 0.4j * X)
    p2 /= np.sqrt(np.sum(np.abs(p2) ** 2) * (Lx / nx) * (Ly / ny))
    out2 = oracle.split2d(p2, V2, -Lx / 2, Lx / 2, -Ly / 2, Ly / 2, -0.05j, 300, method=2,
                          fnes=True)
    kin2, pot2, _, _ = oracle.observables(out2, V2, -Lx / 2, Lx / 2, -Ly / 2, Ly / 2)
    assert kin2 + pot2 == pytest.approx((wx + wy) / 2, abs=1e-6)


def test_device_op_battery_golden_slots():
    """The compiled reference's Device-op battery (oracle/device_op_battery.h)
    holds what its slots name: spot checks of ops whose result is a plain
    formula, so a shifted slot or input cannot pass for the GPU comparison."""
    g = np.load(GOLD)
    a, b, r = g["ops_a"], g["ops_b"], g["ops_r"]
    A, S = g["ops_arrays"], g["ops_scalars"]
    assert np.array_equal(A[0], np.zeros_like(a)) and np.array_equal(A[1], np.ones_like(a))
    assert np.allclose(A[2], np.exp((0.3 - 0.7j) * a), rtol=1e-15, atol=0)
    assert np.array_equal(A[5], a)
    assert np.array_equal(A[10], a - b)
    assert np.allclose(A[11], a * b, rtol=1e-15, atol=1e-16)
    assert np.array_equal(A[21], -a[::-1])
    assert np.array_equal(A[36], r.astype(np.complex128))
    expect_set = a.copy()
    expect_set[3] = 5 - 6j
    assert np.array_equal(A[39], expect_set)
    assert S[2] == a[4] and S[3] == r[5]
    assert S[6] == pytest.approx(np.sum(a.real), rel=1e-14)
    assert S[10] == pytest.approx(np.sum(np.abs(a) ** 2), rel=1e-14)
    assert S[8] == pytest.approx(np.sum(np.abs(a - b) ** 2), rel=1e-14)
