"""Device ops (the Device interface, src/Device.h:53-173) on the GPU against
numpy restatements of the reference functors (src/DeviceTemplate.h)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _rand(n, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def fold(i, n):
    return np.where(i > n // 2, n - i, i)


@pytest.fixture(scope="module")
def N():
    from qusim_amd import _native
    _native.init()
    return _native


_KEEP = []


def up(N, a):
    """Upload; the buffer is kept alive for the whole module (a temporary's
    pointer must not outlive its DeviceArray)."""
    d = N.DeviceArray.from_host(a)
    _KEEP.append(d)
    return d


@pytest.mark.parametrize("n", [0, 1, 1000, 1 << 20])
def test_reductions(N, n):
    a = _rand(n, 1)
    b = _rand(n, 2)
    V = np.random.default_rng(3).standard_normal(n)
    if n == 0:
        da = N.DeviceArray((1,), np.complex128)
    else:
        da, db, dv = up(N, a), up(N, b), up(N, V)
    assert N.reduce1("qsim_hip_norm2", da.ptr, n) == pytest.approx(np.sum(abs(a) ** 2),
                                                                    rel=1e-14, abs=0)
    if n == 0:
        return
    assert N.reduce1("qsim_hip_minus_norm2", da.ptr, db.ptr, n) == pytest.approx(
        np.sum(abs(a - b) ** 2), rel=1e-14)
    assert N.reduce1("qsim_hip_max_abs", da.ptr, n) == pytest.approx(np.max(abs(a)), rel=1e-15)
    assert N.reduce1("qsim_hip_min_abs", da.ptr, n) == pytest.approx(np.min(abs(a)), rel=1e-15)
    assert N.reduce1("qsim_hip_sum_imag", da.ptr, n) == pytest.approx(np.sum(a.imag),
                                                                       rel=1e-12, abs=1e-12)
    i = np.arange(n, dtype=np.float64)
    assert N.reduce1("qsim_hip_abs2_idx", da.ptr, n) == pytest.approx(np.sum(abs(a) ** 2 * i),
                                                                       rel=1e-14)
    assert N.reduce1("qsim_hip_abs2_k1d", da.ptr, 0.3, n) == pytest.approx(
        np.sum(abs(a) ** 2 * fold(i, n) ** 2 * 0.3), rel=1e-14)
    assert N.reduce1("qsim_hip_abs2_mul_real", da.ptr, dv.ptr, n) == pytest.approx(
        np.sum(abs(a) ** 2 * V), rel=1e-12, abs=1e-12)


def test_compensated_norm_exact_on_cancelling_input(N):
    """A sum whose plain FP64 evaluation loses every digit: compensated gives it."""
    n = 1 << 16
    a = np.zeros(n, np.complex128)
    a[0] = 1e8
    a[1:] = 1e-4
    got = N.reduce1("qsim_hip_norm2", up(N, a).ptr, n)
    exact = 1e16 + (n - 1) * 1e-8
    assert abs(got - exact) <= 2.0  # one ulp of 1e16


def test_pointwise(N):
    n = 4099
    a, b = _rand(n, 4), _rand(n, 5)
    V = np.random.default_rng(6).standard_normal(n)
    alpha = 0.3 - 0.7j
    d = up(N, a)
    N.call("qsim_hip_mul_exp_real", d.ptr, N.C128(alpha), up(N, V).ptr, n, None)
    assert rel_l2(d.download(), a * np.exp(alpha * V)) < 1e-14
    d = up(N, a)
    N.call("qsim_hip_mul", d.ptr, up(N, b).ptr, n, None)
    assert rel_l2(d.download(), a * b) < 1e-15
    d = up(N, a)
    N.call("qsim_hip_mul_exp_k1d", d.ptr, N.C128(-0.01j), n, None)
    k = fold(np.arange(n), n).astype(float)
    assert rel_l2(d.download(), a * np.exp(-0.01j * (k * k))) < 1e-13
    nx, ny = 37, 111
    a2 = _rand(nx * ny, 7)
    d = up(N, a2)
    N.call("qsim_hip_mul_exp_k2d", d.ptr, N.C128(-0.2j), 0.01, nx, 0.03, ny, None)
    I, J = np.meshgrid(fold(np.arange(nx), nx), fold(np.arange(ny), ny), indexing="ij")
    want = a2 * np.exp(-0.2j * (I.ravel() ** 2 * 0.01 + J.ravel() ** 2 * 0.03))
    assert rel_l2(d.download(), want) < 1e-13


def test_born_ops(N):
    n = 3001
    dp, p0, V, t = _rand(n, 8), _rand(n, 9), _rand(n, 10) * 0.1, _rand(n, 11)
    eps = 0.4
    out = N.DeviceArray((n,), np.complex128)
    N.call("qsim_hip_add3", out.ptr, up(N, dp).ptr, up(N, p0).ptr, up(N, V).ptr, eps, n, None)
    assert rel_l2(out.download(), (V.real + 1j * (V.imag + eps)) * dp + V.real * p0) < 1e-15
    for name, gam in [("qsim_hip_vellekoop", 1 - 1j * V / eps),
                      ("qsim_hip_hao1", 1 - 1j * np.conj(V) / eps),
                      ("qsim_hip_hao2", 2 / (1 + 1j * V / eps))]:
        d = up(N, dp)
        N.call(name, d.ptr, up(N, t).ptr, up(N, V).ptr, 0.7, eps, n, None)
        g = gam * 0.7
        assert rel_l2(d.download(), (1 - g) * dp + g * t) < 1e-14, name
    nx, ny, nz = 6, 10, 14
    m = nx * ny * nz
    a = _rand(m, 12)
    d = up(N, a)
    N.call("qsim_hip_g03d", d.ptr, 2.0, 0.3, 0.1, 0.2, 0.05, nx, ny, nz, None)
    I, J, K = np.meshgrid(fold(np.arange(nx), nx), fold(np.arange(ny), ny),
                          fold(np.arange(nz), nz), indexing="ij")
    T = I ** 2 * 0.1 + J ** 2 * 0.2 + K ** 2 * 0.05
    assert rel_l2(d.download(), a / (2.0 - T.ravel() + 0.3j)) < 1e-15


def test_xsection3d_reduction(N):
    nx, ny, nz = 8, 6, 10
    n = nx * ny * nz
    p0, p, V = _rand(n, 13), _rand(n, 14), _rand(n, 15)
    x0, y0, z0, dx, dy, dz = -2.0, -1.5, -3.0, 0.5, 0.5, 0.6
    kx, ky, kz = 0.3, -0.2, 1.1
    got = N.reduce_c("qsim_hip_xsection3d", up(N, p0).ptr, up(N, p).ptr, up(N, V).ptr, kx, ky,
                     kz, x0, y0, z0, dx, dy, dz, nx, ny, nz)
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ph = (x0 + I * dx) * kx + (y0 + J * dy) * ky + (z0 + K * dz) * kz
    want = np.sum((p0 + p) * V.real * np.exp(-1j * ph.ravel()))
    assert abs(got - want) / abs(want) < 1e-13


def test_failed_allocation_raises_and_is_not_reported_again(N):
    """An allocation beyond HBM raises (src/DeviceTemplate.h:327-334 throws
    too), and the runtime's stored copy of that error does not reappear as
    the failure of the next kernel launch."""
    with pytest.raises(RuntimeError, match="can't alloc"):
        N.DeviceArray((1 << 41,), np.uint8)   # 2 TiB
    a = up(N, np.ones(1000, np.complex128))
    assert N.reduce1("qsim_hip_norm2", a.ptr, 1000) == 1000.0
