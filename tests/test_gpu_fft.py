"""GPU FFT plans (FourierTransform{1,2,3}D plug-in) against the FP64 oracle.

Sizes include the reference's own test sizes 1234, 123x234, 12x23x34
(applications/QuSimTest/TestFFT.cpp:89-100), every instantiated power of two,
odd primes and length-1 axes."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

# (512, 1024) and (1024, 256): 8-line column tiles of 512 and 1024 points
# (exchanges keyed by the line index) beside 1024- and 256-point rows (the
# swizzled and the padded row placements)
SHAPES = [(1,), (2,), (7,), (17,), (1234,), (16,), (32,), (64,), (128,), (256,), (512,),
          (1024,), (2048,), (4096,), (3, 5), (123, 234), (64, 48), (1024, 1024), (2048, 64),
          (512, 1024), (1024, 256),
          (12, 23, 34), (32, 32, 32), (128, 128, 128), (1, 16, 1), (5, 1, 7)]


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("inverse", [False, True])
def test_fft_fp64_matches_oracle(shape, inverse, oracle):
    from qusim_amd import fftn
    x = _rand(shape, hash(shape) % 1000)
    got = fftn(x, inverse=inverse, precision=64)
    ref = oracle.fft(x, inverse=inverse)
    assert rel_l2(got, ref) < 1e-13


@pytest.mark.parametrize("shape", [(1234,), (1024,), (8192,), (123, 234), (2048, 2048),
                                   (512, 1024), (12, 23, 34), (128, 128, 128)],
                         ids=lambda s: "x".join(map(str, s)))
def test_fft_fp32_matches_oracle(shape, oracle):
    from qusim_amd import fftn
    x = _rand(shape, 3)
    got = fftn(x.astype(np.complex64), precision=32)
    ref = oracle.fft(x.astype(np.complex64).astype(np.complex128))
    assert rel_l2(got, ref) < 2e-6


def test_fft_round_trip_is_n_times_identity():
    """ifft(fft(x)) = N x  (src/kissfft.hh:70-81)"""
    from qusim_amd import fftn
    x = _rand((96, 80), 5)
    y = fftn(fftn(x), inverse=True)
    assert rel_l2(y, x * x.size) < 1e-14


def test_fft_in_place():
    from qusim_amd import FourierTransform
    from qusim_amd._native import DeviceArray
    x = _rand((64, 1234), 6)
    d = DeviceArray.from_host(x)
    FourierTransform(x.shape).transform(d, d)
    assert rel_l2(d.download(), np.fft.fftn(x)) < 1e-13


@pytest.mark.parametrize("shape,prec", [((8192,), 64), ((16384,), 64), ((65536,), 64),
                                        ((3, 16384), 64), ((16384,), 32), ((32768,), 32)],
                         ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else str(v))
@pytest.mark.parametrize("inverse", [False, True])
def test_fft_long_axis_four_step(shape, prec, inverse, oracle):
    """Power-of-two innermost axes beyond one workgroup's LDS run as two
    passes (n = n1 * n2 with a twiddle between and a transposed store)."""
    from qusim_amd import fftn
    x = _rand(shape, 11)
    if prec == 32:
        x = x.astype(np.complex64)
        got = fftn(x, inverse=inverse, precision=32)
        ref = oracle.fft(x.astype(np.complex128), inverse=inverse)
        assert rel_l2(got, ref) < 2e-6
    else:
        got = fftn(x, inverse=inverse, precision=64)
        ref = oracle.fft(x, inverse=inverse)
        assert rel_l2(got, ref) < 1e-13


def test_fft_long_axis_delta_and_round_trip():
    """A shifted delta transforms to one exponential exactly: pins the
    four-step twiddle index and the transposed store position by position."""
    from qusim_amd import fftn
    n, m0 = 16384, 37
    x = np.zeros(n, np.complex128)
    x[m0] = 1.0
    got = fftn(x)
    k = np.arange(n)
    ref = np.exp(-2j * np.pi * ((k * m0) % n) / n)
    assert np.max(np.abs(got - ref)) < 1e-13
    y = fftn(got, inverse=True)
    assert np.max(np.abs(y - n * x)) < 1e-9
