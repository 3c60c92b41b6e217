"""Device-side input generation against the host formulas of the benchmark
workloads (qusim_amd/workloads.py, numpy): Gaussian-bump potentials with the
imaginary absorbing layer, and normalised Gaussian packets.

The device evaluates the same expressions in the same order; only exp / sin /
cos round differently (each within an ulp or so), so the bar is 1e-14
relative (max norm for potentials, L2 for packets)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _arr(v, dt=ctypes.c_double):
    return (dt * len(v))(*v)


def gen_potential(dim, n, lo, hi, bumps, absorb_width=0.0, absorb_strength=0.0, out="split"):
    from qusim_amd import _native as N
    total = int(np.prod(n))
    flat = [x for b in bumps for x in b]
    nb = _arr([int(x) for x in n], ctypes.c_size_t)
    if out == "split":
        re = N.DeviceArray(total, np.float64)
        im = N.DeviceArray(total, np.float64)
        N.call("qsim_hip_gen_potential", dim, nb, _arr(lo), _arr(hi), _arr(flat) if flat else None,
               len(bumps), absorb_width, absorb_strength, re.ptr, im.ptr, None, None)
        return (re.download() + 1j * im.download()).reshape(n)
    cv = N.DeviceArray(total, np.complex128)
    N.call("qsim_hip_gen_potential", dim, nb, _arr(lo), _arr(hi), _arr(flat) if flat else None,
           len(bumps), absorb_width, absorb_strength, None, None, cv.ptr, None)
    return cv.download().reshape(n)


def gen_packet(dim, n, lo, hi, center, k, sigma, normalise=True):
    from qusim_amd import _native as N
    total = int(np.prod(n))
    psi = N.DeviceArray(total, np.complex128)
    N.call("qsim_hip_gen_packet", dim, _arr([int(x) for x in n], ctypes.c_size_t), _arr(lo),
           _arr(hi), _arr(center), _arr(k), sigma, int(normalise), psi.ptr, None)
    return psi.download().reshape(n)


def maxrel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_c2_potential_and_packet():
    from conftest import rel_l2
    from qusim_amd.workloads import bumps_2d, c2_problem
    p = c2_problem()
    bumps = [(h, w, cx, cy) for cx, cy, h, w in bumps_2d()]
    V = gen_potential(2, (1024, 1024), (-150.0, -150.0), (150.0, 150.0), bumps, 20.0, 0.1)
    assert maxrel(V.real, p["V"].real) < 1e-14
    assert np.array_equal(V.imag, p["V"].imag)  # the layer has no transcendental
    psi = gen_packet(2, (1024, 1024), (-150.0, -150.0), (150.0, 150.0), (-60.0, 0.0),
                     (0.8, 0.0), 8.0)
    assert rel_l2(psi, p["psi"]) < 1e-14


def test_c1_barrier_and_packet():
    from conftest import rel_l2
    from qusim_amd.workloads import c1_problem
    p = c1_problem()
    V = gen_potential(1, (2048,), (-100.0,), (100.0,), [(0.6, 1.0, 0.0)])
    assert maxrel(V.real, p["V"]) < 1e-14 and not V.imag.any()
    psi = gen_packet(1, (2048,), (-100.0,), (100.0,), (-30.0,), (1.0,), 5.0)
    assert rel_l2(psi, p["psi"]) < 1e-14


def test_c4_potential_complex_output():
    from qusim_amd.workloads import c4_potential, gaussians_3d
    n = 64
    ref = c4_potential(n=n)
    bumps = [(a, w, c[0], c[1], c[2]) for a, w, c in gaussians_3d()]
    V = gen_potential(3, (n, n, n), (-32.0,) * 3, (32.0,) * 3, bumps, out="c128")
    assert maxrel(V, ref) < 1e-14


def test_unnormalised_packet_and_errors():
    from qusim_amd import _native as N
    x = -5 + 10.0 / 40 * np.arange(40)
    psi = gen_packet(1, (40,), (-5.0,), (5.0,), (0.5,), (2.0,), 1.5, normalise=False)
    ref = np.exp(1j * 2.0 * x - (x - 0.5) ** 2 / (4 * 1.5 ** 2))
    assert np.max(np.abs(psi - ref)) < 1e-15
    with pytest.raises(RuntimeError):
        gen_potential(4, (2, 2, 2, 2), (0.0,) * 4, (1.0,) * 4, [])
    with pytest.raises(RuntimeError):
        gen_potential(1, (8,), (0.0,), (1.0,), [(1.0, 0.0, 0.5)])   # width 0
    with pytest.raises(RuntimeError):
        gen_potential(1, (8,), (1.0,), (1.0,), [])                    # empty box
    assert N is not None
