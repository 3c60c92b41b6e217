"""ctypes wrapper of the CPU oracle (liboracle.so) and of the compiled
reference checker (_ref/libqsim_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product path never imports this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_VP = ctypes.c_void_p
_SZ = ctypes.c_size_t
_D = ctypes.c_double
_I = ctypes.c_int

_lib = None
_ref = None


def lib():
    """Load liboracle.so (built by `make -C oracle`)."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (%s missing)" % path)
        L = ctypes.CDLL(path)
        L.orc_fft.argtypes = [_I, _VP, _I, _VP, _VP, _I]
        L.orc_dft_bruteforce.argtypes = [_SZ, _I, _VP, _VP]
        L.orc_norm2.argtypes = [_VP, _SZ]
        L.orc_norm2.restype = _D
        L.orc_split1d_run.argtypes = [_SZ, _D, _D, _D, _D, _D, _D, _I, _I, _VP, _I, _VP,
                                      _SZ, _SZ, _I, _I]
        L.orc_split2d_run.argtypes = [_SZ, _SZ, _D, _D, _D, _D, _D, _D, _D, _D, _I, _VP, _I,
                                      _VP, _SZ, _SZ, _I, _I]
        L.orc_split2d_time.argtypes = [_SZ, _SZ, _D, _D, _D, _D, _D, _D, _D, _VP, _I, _VP,
                                       _SZ, _I, _I]
        L.orc_split2d_time.restype = _D
        L.orc_born_time.argtypes = [_I, _VP, _VP, _VP, _D, _VP, _D, _D, _I, _I, _D, _I, _VP,
                                    _I, _I]
        L.orc_born_time.restype = _D
        L.orc_norm2_range_1d.argtypes = [_VP, _SZ, _D, _D, _D, _D]
        L.orc_norm2_range_1d.restype = _D
        L.orc_observables.argtypes = [_I, _SZ, _SZ, _D, _D, _D, _D, _D, _D, _I, _VP, _VP, _VP]
        L.orc_born_run.argtypes = [_I, _VP, _VP, _VP, _D, _VP, _D, _D, _I, _I, _D, _I, _VP,
                                   _VP, _VP, _VP, _VP, _I, _I]
        L.orc_plane_wave.argtypes = [_I, _VP, _VP, _VP, _D, _VP, _D, _D, _VP]
        L.orc_xsection2d.argtypes = [_VP, _VP, _VP, _D, _D, _D, _VP, _VP, _VP, _D, _D]
        L.orc_xsection2d.restype = _D
        L.orc_xsection3d.argtypes = [_VP, _VP, _VP, _D, _D, _D, _VP, _VP, _VP, _D, _D, _D, _VP]
        L.orc_xsection3d.restype = _D
        L.orc_total_xsection2d.argtypes = [_VP, _VP, _VP, _D, _D, _D, _VP, _VP, _VP,
                                           ctypes.c_longlong]
        L.orc_total_xsection2d.restype = _D
        L.orc_total_xsection3d.argtypes = [_VP, _VP, _VP, _D, _D, _D, _VP, _VP, _VP,
                                           ctypes.c_longlong, ctypes.c_longlong]
        L.orc_total_xsection3d.restype = _D
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libqsim_ref.so"))


def ref():
    """Load the compiled-reference checker (oracle/_ref/libqsim_ref.so)."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libqsim_ref.so")
        if not os.path.exists(path):
            raise RuntimeError("compiled reference not built (%s missing)" % path)
        R = ctypes.CDLL(path)
        R.ref_fft.argtypes = [_I, _VP, _I, _VP, _VP]
        R.ref_split1d_run.argtypes = [_SZ, _D, _D, _D, _D, _D, _D, _I, _I, _VP, _VP, _SZ, _SZ]
        R.ref_split2d_run.argtypes = [_SZ, _SZ, _D, _D, _D, _D, _D, _D, _D, _D, _I, _VP, _VP,
                                      _SZ, _SZ]
        R.ref_born_run.argtypes = [_I, _VP, _VP, _VP, _D, _VP, _D, _D, _I, _I, _D, _I, _VP,
                                   _VP, _VP, _VP, _VP, _I]
        R.ref_xsection3d.argtypes = [_VP, _VP, _VP, _D, _D, _D, _VP, _VP, _VP, _D, _D, _D]
        R.ref_xsection3d.restype = _D
        R.ref_split1d_cv_run.argtypes = [_SZ, _D, _D, _D, _D, _D, _D, _I, _I, _VP, _VP, _SZ,
                                         _SZ, _I]
        R.ref_split2d_cv_run.argtypes = [_SZ, _SZ, _D, _D, _D, _D, _D, _D, _D, _D, _I, _VP,
                                         _VP, _SZ, _SZ, _I]
        R.ref_observables.argtypes = [_I, _SZ, _SZ, _D, _D, _D, _D, _D, _D, _I, _VP, _VP, _VP]
        R.ref_device_ops.argtypes = [_SZ, _SZ, _SZ, _VP, _VP, _VP, _VP, _VP]
        _ref = R
    return _ref


def _c(a, dtype=np.complex128):
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a):
    return a.ctypes.data_as(_VP)


def _szs(seq):
    return (ctypes.c_size_t * len(seq))(*seq)


def _ds(seq):
    return (ctypes.c_double * len(seq))(*seq)


# ---------------------------------------------------------------- oracle
def fft(x, inverse=False, threads=1):
    x = _c(x)
    out = np.empty_like(x)
    rc = lib().orc_fft(x.ndim, _szs(x.shape), int(bool(inverse)), _p(x), _p(out), threads)
    assert rc == 0
    return out


def dft_bruteforce(x, inverse=False):
    x = _c(x)
    out = np.empty_like(x)
    lib().orc_dft_bruteforce(x.size, int(bool(inverse)), _p(x), _p(out))
    return out


def split1d(psi, V, x0, x1, dt, nsteps, method=2, boundary=0, mass=1.0, hbar=1.0,
            complex_v=False, batch=1, fnes=False, threads=1):
    psi = _c(psi).copy()
    V = _c(V)
    dt = complex(dt)
    rc = lib().orc_split1d_run(psi.size, x0, x1, dt.real, dt.imag, mass, hbar, method,
                               boundary, _p(V), int(complex_v), _p(psi), nsteps, batch,
                               int(fnes), threads)
    assert rc == 0
    return psi


def split2d(psi, V, x0, x1, y0, y1, dt, nsteps, method=2, mass=1.0, hbar=1.0,
            complex_v=False, batch=1, fnes=False, threads=1):
    psi = _c(psi).copy()
    V = _c(V)
    dt = complex(dt)
    nx, ny = psi.shape
    rc = lib().orc_split2d_run(nx, ny, x0, x1, y0, y1, dt.real, dt.imag, mass, hbar, method,
                               _p(V), int(complex_v), _p(psi), nsteps, batch, int(fnes),
                               threads)
    assert rc == 0
    return psi


def split2d_time(psi, V, x0, x1, y0, y1, dt, nsteps, complex_v=True, threads=1,
                 fft_threads=None):
    """Seconds spent in `nsteps` O2 steps of the oracle loop (set-up excluded);
    `threads` for the pointwise ops, `fft_threads` (default: threads) for the
    FFT lines."""
    psi = _c(psi).copy()
    V = _c(V)
    nx, ny = psi.shape
    return lib().orc_split2d_time(nx, ny, x0, x1, y0, y1, dt, 1.0, 1.0, _p(V), int(complex_v),
                                  _p(psi), nsteps, threads,
                                  threads if fft_threads is None else fft_threads)


def born_time(V, lo, hi, energy, direction, iters, preconditional=True, preconditioner=1,
              slow=1.0, mass=1.0, hbar=1.0, threads=1, fft_threads=None):
    """Seconds of `iters` oracle Born iterations (the Compute() loop only)."""
    V = _c(V)
    s = lib().orc_born_time(V.ndim, _szs(V.shape), _ds(lo), _ds(hi), energy, _ds(direction),
                            mass, hbar, int(preconditional), preconditioner, slow, iters, _p(V),
                            threads, threads if fft_threads is None else fft_threads)
    assert s >= 0
    return s


def norm2(a):
    a = _c(a)
    return lib().orc_norm2(_p(a), a.size)


def norm2_range_1d(psi, x0, x1, a, b):
    psi = _c(psi)
    return lib().orc_norm2_range_1d(_p(psi), psi.size, x0, x1, a, b)


def observables(psi, V, x0, x1, y0=0.0, y1=1.0, mass=1.0, hbar=1.0, boundary=0):
    """(KinEn, PotEn, Norm2, Xavg) of the evolver state psi (1D or 2D) with
    the real potential V; Xavg is 0 in 2D."""
    psi = _c(psi)
    V = np.ascontiguousarray(np.real(V), dtype=np.float64)
    out = (ctypes.c_double * 4)()
    nx = psi.shape[0]
    ny = psi.shape[1] if psi.ndim == 2 else 1
    rc = lib().orc_observables(psi.ndim, nx, ny, x0, x1, y0, y1, mass, hbar, boundary,
                               _p(psi), _p(V), out)
    assert rc == 0
    return tuple(out)


def born(V, lo, hi, energy, direction, order, ncalls=1, preconditional=True,
         preconditioner=1, slow=1.0, epsilon=0.0, mass=1.0, hbar=1.0, dpsi=None, threads=1):
    """Returns (dpsi, v_eff, epsilon_used, delta_norm)."""
    V = _c(V)
    dim = V.ndim
    dp = np.zeros_like(V) if dpsi is None else _c(dpsi).copy()
    veff = np.empty_like(V)
    eps = ctypes.c_double(epsilon)
    dn = ctypes.c_double(0)
    rc = lib().orc_born_run(dim, _szs(V.shape), _ds(lo), _ds(hi), energy, _ds(direction),
                            mass, hbar, int(preconditional), preconditioner, slow, order,
                            _p(V), _p(dp), _p(veff), ctypes.byref(eps), ctypes.byref(dn),
                            ncalls, threads)
    assert rc == 0
    return dp, veff, eps.value, dn.value


def plane_wave(shape, lo, hi, energy, direction, mass=1.0, hbar=1.0):
    out = np.empty(shape, np.complex128)
    lib().orc_plane_wave(len(shape), _szs(shape), _ds(lo), _ds(hi), energy, _ds(direction),
                         mass, hbar, _p(out))
    return out


def xsection2d(psi0, dpsi, V, lo, hi, energy, cosx, cosy, mass=1.0, hbar=1.0):
    psi0, dpsi, V = _c(psi0), _c(dpsi), _c(V)
    return lib().orc_xsection2d(_szs(V.shape), _ds(lo), _ds(hi), energy, mass, hbar,
                                _p(psi0), _p(dpsi), _p(V), cosx, cosy)


def xsection3d(psi0, dpsi, V, lo, hi, energy, cosx, cosy, cosz, mass=1.0, hbar=1.0):
    psi0, dpsi, V = _c(psi0), _c(dpsi), _c(V)
    amp = (ctypes.c_double * 2)()
    x = lib().orc_xsection3d(_szs(V.shape), _ds(lo), _ds(hi), energy, mass, hbar, _p(psi0),
                             _p(dpsi), _p(V), cosx, cosy, cosz, amp)
    return x, complex(amp[0], amp[1])


def total_xsection2d(psi0, dpsi, V, lo, hi, energy, n, mass=1.0, hbar=1.0):
    psi0, dpsi, V = _c(psi0), _c(dpsi), _c(V)
    return lib().orc_total_xsection2d(_szs(V.shape), _ds(lo), _ds(hi), energy, mass, hbar,
                                      _p(psi0), _p(dpsi), _p(V), n)


def total_xsection3d(psi0, dpsi, V, lo, hi, energy, npsi, ntheta, mass=1.0, hbar=1.0):
    psi0, dpsi, V = _c(psi0), _c(dpsi), _c(V)
    return lib().orc_total_xsection3d(_szs(V.shape), _ds(lo), _ds(hi), energy, mass, hbar,
                                      _p(psi0), _p(dpsi), _p(V), npsi, ntheta)


# ---------------------------------------------------------------- compiled reference
def ref_fft(x, inverse=False):
    x = _c(x)
    out = np.empty_like(x)
    rc = ref().ref_fft(x.ndim, _szs(x.shape), int(bool(inverse)), _p(x), _p(out))
    assert rc == 0
    return out


def ref_split1d(psi, Vreal, x0, x1, dt, nsteps, method=2, boundary=0, mass=1.0, hbar=1.0,
                batch=1):
    psi = _c(psi).copy()
    V = np.ascontiguousarray(Vreal, dtype=np.float64)
    dt = complex(dt)
    ref().ref_split1d_run(psi.size, x0, x1, dt.real, dt.imag, mass, hbar, method, boundary,
                          _p(V), _p(psi), nsteps, batch)
    return psi


def ref_split2d(psi, Vreal, x0, x1, y0, y1, dt, nsteps, method=2, mass=1.0, hbar=1.0,
                batch=1):
    psi = _c(psi).copy()
    V = np.ascontiguousarray(Vreal, dtype=np.float64)
    dt = complex(dt)
    nx, ny = psi.shape
    ref().ref_split2d_run(nx, ny, x0, x1, y0, y1, dt.real, dt.imag, mass, hbar, method, _p(V),
                          _p(psi), nsteps, batch)
    return psi


def ref_split1d_cv(psi, V, x0, x1, dt, nsteps, method=2, boundary=0, mass=1.0, hbar=1.0,
                   batch=1, fnes=False):
    """The reference's 1D split step with a complex potential (its complex
    MulExp device op) and the optional FNES renormalisation of Step()."""
    psi = _c(psi).copy()
    V = _c(V)
    dt = complex(dt)
    ref().ref_split1d_cv_run(psi.size, x0, x1, dt.real, dt.imag, mass, hbar, method, boundary,
                             _p(V), _p(psi), nsteps, batch, int(fnes))
    return psi


def ref_split2d_cv(psi, V, x0, x1, y0, y1, dt, nsteps, method=2, mass=1.0, hbar=1.0, batch=1,
                   fnes=False):
    psi = _c(psi).copy()
    V = _c(V)
    dt = complex(dt)
    nx, ny = psi.shape
    ref().ref_split2d_cv_run(nx, ny, x0, x1, y0, y1, dt.real, dt.imag, mass, hbar, method,
                             _p(V), _p(psi), nsteps, batch, int(fnes))
    return psi


def ref_observables(psi, V, x0, x1, y0=0.0, y1=1.0, mass=1.0, hbar=1.0, boundary=0):
    psi = _c(psi)
    V = np.ascontiguousarray(np.real(V), dtype=np.float64)
    out = (ctypes.c_double * 4)()
    nx = psi.shape[0]
    ny = psi.shape[1] if psi.ndim == 2 else 1
    rc = ref().ref_observables(psi.ndim, nx, ny, x0, x1, y0, y1, mass, hbar, boundary,
                               _p(psi), _p(V), out)
    assert rc == 0
    return tuple(out)


def ref_born(V, lo, hi, energy, direction, order, ncalls=1, preconditional=True,
             preconditioner=1, slow=1.0, epsilon=0.0, mass=1.0, hbar=1.0):
    V = _c(V)
    dp = np.zeros_like(V)
    veff = np.empty_like(V)
    eps = ctypes.c_double(epsilon)
    dn = ctypes.c_double(0)
    rc = ref().ref_born_run(V.ndim, _szs(V.shape), _ds(lo), _ds(hi), energy, _ds(direction),
                            mass, hbar, int(preconditional), preconditioner, slow, order,
                            _p(V), _p(dp), _p(veff), ctypes.byref(eps), ctypes.byref(dn),
                            ncalls)
    assert rc == 0
    return dp, veff, eps.value, dn.value


# Slots of the device-op battery (oracle/device_op_battery.h)
OPS_ARRAYS = 41
OPS_SCALARS = 18


def ref_device_ops(dims, a, b, c, r):
    """Every Device op of the compiled reference's CpuSeq device on the
    battery's inputs: (arrays [OPS_ARRAYS, n], scalars [OPS_SCALARS])."""
    nx, ny, nz = dims
    n = nx * ny * nz
    a, b, c = _c(a), _c(b), _c(c)
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros(n * OPS_ARRAYS + OPS_SCALARS, np.complex128)
    rc = ref().ref_device_ops(nx, ny, nz, _p(a), _p(b), _p(c), _p(r), _p(out))
    assert rc == 0
    return out[:n * OPS_ARRAYS].reshape(OPS_ARRAYS, n), out[n * OPS_ARRAYS:]


def ref_xsection3d(psi0, dpsi, V, lo, hi, energy, cosx, cosy, cosz, mass=1.0, hbar=1.0):
    psi0, dpsi, V = _c(psi0), _c(dpsi), _c(V)
    return ref().ref_xsection3d(_szs(V.shape), _ds(lo), _ds(hi), energy, mass, hbar, _p(psi0),
                                _p(dpsi), _p(V), cosx, cosy, cosz)
