"""Host-side mirror of the reference public API (include/QuSim.h) on the HIP
device: Options, Evolver1D / Evolver2D, QuPerturbation2D / QuPerturbation3D.

Names, argument order and meaning follow the reference so callers port
line by line; errors raise as the reference throws.  All numerical work runs
through libqsim_hip.so (include/qsim_hip.h); only the per-point evaluation of
the user's psi/V functions at init happens on the host, as in the reference
(src/EvolverImpl2D.cpp:40-99)."""
import ctypes
import enum
import math

import numpy as np

from . import _native as N
from . import snapshot
from .fft import FourierTransform


class BoundaryCondition(enum.Enum):
    """include/QuSim.h:467-470"""
    Period = 0
    InfiniteWall = 1


class SolverMethod(enum.Enum):
    """include/QuSim.h:475-494 (the subset this device runs)."""
    Unknown = 0
    SplittingMethodO2 = 1
    SplittingMethodO4 = 2
    BornSerise = 9


class Options:
    """String-keyed option store with the reference's fluent setters
    (include/QuSim.h:285-458) plus Hip() and Precision()."""

    def __init__(self, other=None):
        self._d = dict(other._d) if other is not None else {}

    # ---- generic access (src/OptionsImpl.cpp)
    def SetInt(self, k, v):
        self._d[k] = int(v)
        return self

    def SetReal(self, k, v):
        self._d[k] = float(v)
        return self

    def SetComplex(self, k, v):
        self._d[k] = complex(v)
        return self

    def SetBool(self, k, v):
        self._d[k] = bool(v)
        return self

    def SetString(self, k, v):
        self._d[k] = str(v)
        return self

    def Get(self, k, default=None):
        return self._d.get(k, default)

    def Contains(self, k):
        return k in self._d

    # ---- reference setters
    def Order(self, n):
        return self.SetInt("order", n)

    def VellekoopPreconditioner(self):
        return self.SetString("preconditioner", "Vellekoop")

    def Hao1Preconditioner(self):
        return self.SetString("preconditioner", "Hao1")

    def Hao2Preconditioner(self):
        return self.SetString("preconditioner", "Hao2")

    def BornIdentityPreconditioner(self):
        return self.SetString("preconditioner", "BornIndentity")

    def Preconditional(self, v):
        return self.SetBool("preconditional", v)

    def Slow(self, v):
        return self.SetReal("slow", v)

    def Batch(self, n):
        return self.SetInt("batch", n)

    def ComplexPotential(self, cp=True):
        return self.SetBool("complex_potential", cp)

    def Cuda(self):
        return self.SetString("device", "GPU_CUDA")

    def CpuSeq(self):
        # the reference writes "GPU_SEQ" here (include/QuSim.h:354); fixed
        return self.SetString("device", "CPU_SEQ")

    def CpuPar(self):
        return self.SetString("device", "CPU_PAR")

    def CpuParVec(self):
        return self.SetString("device", "CPU_PAR_VEC")

    def CpuVec(self):
        return self.SetString("device", "CPU_VEC")

    # ---- additions of this device
    def Hip(self):
        return self.SetString("device", "GPU_HIP")

    def Precision(self, p):
        """"fp64" (complex128, default) or "mixed32" (FP32 data, FP64 phases and
        reductions).  The reference's dormant CudaPrecisionSingle() maps to
        "mixed32"."""
        if p not in ("fp64", "mixed32"):
            raise ValueError("invalid precision: " + str(p))
        return self.SetString("precision", p)

    def CudaPrecisionSingle(self):
        return self.Precision("mixed32")


def _device_type(opts):
    """GetDeviceType (src/DeviceType.h:6-24) with GPU_HIP added.  This package
    is the HIP device, so it is also the default; the reference's CPU devices
    and CUDA are not provided here."""
    dev = opts.Get("device", "GPU_HIP")
    if dev == "GPU_HIP":
        return dev
    if dev in ("CPU_SEQ", "CPU_PAR", "CPU_PAR_VEC", "CPU_VEC", "GPU_CUDA"):
        raise ValueError("device %s is not provided by qusim_amd (use Options().Hip())" % dev)
    raise ValueError("invalid device: " + str(dev))


def _precision(opts):
    p = opts.Get("precision", "fp64")
    return 64 if p == "fp64" else 32


def _eval1(f, xs):
    """Evaluate a user function on the grid: vectorised when it accepts arrays,
    else point by point (the reference calls it per point)."""
    try:
        out = np.asarray(f(xs), dtype=np.complex128)
        if out.shape == xs.shape:
            return out
    except Exception:
        pass
    return np.array([complex(f(float(x))) for x in xs], dtype=np.complex128)


def _eval2(f, X, Y):
    try:
        out = np.asarray(f(X, Y), dtype=np.complex128)
        if out.shape == X.shape:
            return out
    except Exception:
        pass
    return np.array([complex(f(float(x), float(y))) for x, y in zip(X.ravel(), Y.ravel())],
                    dtype=np.complex128).reshape(X.shape)


def _eval3(f, X, Y, Z):
    try:
        out = np.asarray(f(X, Y, Z), dtype=np.complex128)
        if out.shape == X.shape:
            return out
    except Exception:
        pass
    return np.array([complex(f(float(x), float(y), float(z)))
                     for x, y, z in zip(X.ravel(), Y.ravel(), Z.ravel())],
                    dtype=np.complex128).reshape(X.shape)


# ------------------------------------------------------------------ Evolver
class Evolver:
    """Common part of Evolver1D/2D (src/EvolverImpl.cpp)."""

    def _setup(self, psi_host, V, force_normalization, dt, fnes, boundary, solver, mass, hbar,
               dv, opts, dim, nx, ny, x0, x1, y0, y1):
        _device_type(opts)
        if solver not in (SolverMethod.SplittingMethodO2, SolverMethod.SplittingMethodO4):
            raise RuntimeError("unsupported solver")
        self._opts = Options(opts)
        self._dt = complex(dt)
        self._mass = float(mass)
        self._hbar = float(hbar)
        self._dv = float(dv)
        self._fnes = bool(fnes)
        self._batch = int(opts.Get("batch", 1))
        self._method = solver
        self._boundary = boundary
        self._prec = _precision(opts)
        self.fStep = 0
        psi_host = np.ascontiguousarray(psi_host, dtype=np.complex128)
        if force_normalization:  # src/EvolverImpl2D.cpp:87-92
            psi_host = psi_host / math.sqrt(float(np.sum(np.abs(psi_host) ** 2)) * self._dv)
        self._complex_v = bool(opts.Get("complex_potential", False))
        self._V_host = V if self._complex_v else V.real.astype(np.complex128)
        self.psi = N.DeviceArray.from_host(psi_host)
        self._vre = N.DeviceArray.from_host(np.ascontiguousarray(V.real, dtype=np.float64))
        self._vim = (N.DeviceArray.from_host(np.ascontiguousarray(V.imag, dtype=np.float64))
                     if self._complex_v else None)
        d = N.SplitDesc()
        d.dim, d.nx, d.ny = dim, nx, ny
        d.x0, d.x1, d.y0, d.y1 = x0, x1, y0, y1
        d.dt_re, d.dt_im = self._dt.real, self._dt.imag
        d.mass, d.hbar = self._mass, self._hbar
        d.method = 2 if solver == SolverMethod.SplittingMethodO2 else 4
        d.boundary = 0 if boundary == BoundaryCondition.Period else 1
        d.precision = self._prec
        p = ctypes.c_void_p()
        N.call("qsim_hip_split_create", ctypes.byref(p), ctypes.byref(d), self._vre.ptr,
               self._vim.ptr if self._vim is not None else None, self.psi.ptr)
        self._job = p.value
        self._n = self.psi.size

    def step(self):
        """QuEvolverImpl::Step (src/EvolverImpl.cpp:56-68): `batch` split steps,
        then the optional renormalisation.  psi stays on the device; GetPsi()
        copies it out."""
        N.call("qsim_hip_split_steps", self._job, self._batch, None)
        self.fStep += self._batch
        if self._fnes:
            sc = 1.0 / math.sqrt(self.Norm2())
            N.call("qsim_hip_scale", self.psi.ptr, sc, self._n, None)

    def steps(self, n):
        """n step() calls."""
        if not self._fnes:
            N.call("qsim_hip_split_steps", self._job, self._batch * n, None)
            self.fStep += self._batch * n
        else:
            for _ in range(n):
                self.step()

    def Norm2(self):
        return N.reduce1("qsim_hip_norm2", self.psi.ptr, self._n) * self._dv

    def Time(self):
        return (self.fStep * self._dt).real

    def PotEn(self):
        """QuEvolverImpl::CalPotEn (src/EvolverImpl.cpp:45-49)."""
        vre = self._vre
        e = N.reduce1("qsim_hip_abs2_mul_real", self.psi.ptr, vre.ptr, self._n) * self._dv
        return e / self.Norm2()

    def GetMethod(self):
        return self._method

    def _snapshot_header(self):
        d = len(self._shape)
        return {"dim": d, "extent": list(self._extent), "dt": [self._dt.real, self._dt.imag],
                "mass": self._mass, "hbar": self._hbar,
                "method": "O2" if self._method == SolverMethod.SplittingMethodO2 else "O4",
                "boundary": self._boundary.name}

    def SaveSnapshot(self, path):
        """Write psi and the step count to `path` (format in
        qusim_amd/snapshot.py).  No counterpart in the reference."""
        h = self._snapshot_header()
        h.update(step=self.fStep, precision=self._prec)
        snapshot.write(path, h, self.psi.download().reshape(self._shape))

    def LoadSnapshot(self, path):
        """Replace psi and the step count with a snapshot's.  The grid, dt,
        mass, hbar, method and boundary must equal this evolver's (ValueError
        otherwise); the precision may differ, since psi is complex128 either
        way.  The next steps then continue the saved evolution bit for bit."""
        h, psi = snapshot.read(path)
        mine = self._snapshot_header()
        if tuple(psi.shape) != tuple(self._shape):
            raise ValueError("snapshot grid %s, evolver grid %s" % (psi.shape, self._shape))
        for k, v in mine.items():
            if h.get(k) != v:
                raise ValueError("snapshot %s %r, evolver %s %r" % (k, h.get(k), k, v))
        step = h.get("step")
        if not isinstance(step, int) or step < 0:
            raise ValueError("snapshot step %r is not a step count" % (step,))
        self.psi.upload(psi)
        self.fStep = step

    def close(self):
        if getattr(self, "_job", None):
            N.lib().qsim_hip_split_destroy(self._job)
            self._job = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Evolver1D(Evolver):
    def init(self, psi, force_normalization, dt, force_normalization_each_step, v, x0, x1, n,
             b, solver, mass, hbar, opts):
        """Evolver1D::init (include/QuSim.h:561-566, src/EvolverImpl1D.cpp:4-30)."""
        self.fNx = int(n)
        self.fX0, self.fX1 = float(x0), float(x1)
        self.fDx = (self.fX1 - self.fX0) / self.fNx
        xs = self.fDx * np.arange(self.fNx) + self.fX0
        psi_h = psi if isinstance(psi, np.ndarray) else _eval1(psi, xs)
        V = v if isinstance(v, np.ndarray) else _eval1(v, xs)
        self._shape = (self.fNx,)
        self._extent = (self.fX0, self.fX1)
        self._setup(psi_h, np.asarray(V, np.complex128), force_normalization, dt,
                    force_normalization_each_step, b, solver, mass, hbar, self.fDx, opts, 1,
                    self.fNx, 1, self.fX0, self.fX1, 0.0, 1.0)
        return self

    def GetPsi(self):
        return self.psi.download()

    def GetV(self):
        return self._V_host.real.copy()

    def GetN(self):
        return self.fNx

    def GetX0(self):
        return self.fX0

    def GetX1(self):
        return self.fX1

    def Norm2(self, a=None, b=None):
        """Norm2() or Norm2(x0, x1) (src/EvolverImpl1D.cpp:89-101)."""
        if a is None:
            return Evolver.Norm2(self)
        i1 = int((a - self.fX0) / self.fDx)
        i2 = int((b - self.fX0) / self.fDx)
        i1 = max(i1, 0)
        i2 = min(i2, self.fNx)
        if i1 >= i2:
            return 0.0
        return N.reduce1("qsim_hip_norm2", self.psi.ptr + 16 * i1, i2 - i1) * self._dv

    def Xavg(self):
        """src/EvolverImpl1D.cpp:82-87"""
        idx = N.reduce1("qsim_hip_abs2_idx", self.psi.ptr, self._n) / N.reduce1(
            "qsim_hip_norm2", self.psi.ptr, self._n)
        return self.fX0 + self.fDx * idx

    def KinEn(self):
        """SplittingMethod1D::CalKinEn (src/SplittingMethod1D.cpp:94-128).
        Periodic: sum |FFT psi|^2 fold(i)^2 DT / sum |FFT psi|^2 with
        DT = Dk^2 / 2m (the reference leaves hbar out here).  Infinite wall:
        the odd extension to 2N through the inverse 2N transform, then
        sum_{i<N} |b_i|^2 i^2 DT / sum_{i<N} |b_i|^2, DT = (hbar pi/(dx N))^2 / 2m."""
        n = self.fNx
        if self._boundary == BoundaryCondition.Period:
            f = FourierTransform((n,), False)
            tmp = N.DeviceArray((n,), np.complex128)
            f.transform(self.psi, tmp)
            Dk = 2 * math.pi / (self.fDx * n)
            DT = Dk * Dk / (2 * self._mass)
            kin = N.reduce1("qsim_hip_abs2_k1d", tmp.ptr, DT, n)
            return kin / N.reduce1("qsim_hip_norm2", tmp.ptr, n)
        a = N.DeviceArray((2 * n,), np.complex128)
        b = N.DeviceArray((2 * n,), np.complex128)
        zero = N.c128(0.0, 0.0)
        N.call("qsim_hip_copy", a.ptr, self.psi.ptr, n, None)
        N.call("qsim_hip_set_elem", a.ptr, 0, zero, None)
        N.call("qsim_hip_set_elem", a.ptr, n, zero, None)
        N.call("qsim_hip_copy_reverse_minus", a.ptr + 16 * (2 * n - 1), a.ptr + 16, n - 1, None)
        FourierTransform((2 * n,), True).transform(a, b)
        Dk = math.pi / (self.fDx * n)
        DT = (Dk * self._hbar) ** 2 / (2 * self._mass)
        kin = N.reduce1("qsim_hip_abs2_idx2", b.ptr, n) * DT
        return kin / N.reduce1("qsim_hip_norm2", b.ptr, n)


class Evolver2D(Evolver):
    def init(self, psi, force_normalization, dt, force_normalization_each_step, v, x0, x1, nx,
             y0, y1, ny, b, solver, mass, hbar, opts):
        """Evolver2D::init (include/QuSim.h:585-591, src/EvolverImpl2D.cpp:5-38)."""
        if b != BoundaryCondition.Period:
            raise RuntimeError("unsupported boundary condition!")
        self.fNx, self.fNy = int(nx), int(ny)
        self.fX0, self.fY0 = float(x0), float(y0)
        self.fDx = (float(x1) - self.fX0) / self.fNx
        self.fDy = (float(y1) - self.fY0) / self.fNy
        xs = self.fDx * np.arange(self.fNx) + self.fX0
        ys = self.fDy * np.arange(self.fNy) + self.fY0
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        psi_h = psi if isinstance(psi, np.ndarray) else _eval2(psi, X, Y)
        V = v if isinstance(v, np.ndarray) else _eval2(v, X, Y)
        self._shape = (self.fNx, self.fNy)
        self._extent = (self.fX0, float(x1), self.fY0, float(y1))
        self._setup(psi_h, np.asarray(V, np.complex128), force_normalization, dt,
                    force_normalization_each_step, b, solver, mass, hbar, self.fDx * self.fDy,
                    opts, 2, self.fNx, self.fNy, self.fX0, float(x1), self.fY0, float(y1))
        return self

    def GetPsi(self):
        return self.psi.download().reshape(self.fNx, self.fNy)

    def GetV(self):
        return self._V_host.real.reshape(self.fNx, self.fNy).copy()

    def GetNx(self):
        return self.fNx

    def GetNy(self):
        return self.fNy

    def KinEn(self):
        """SplittingMethod2D::CalKinEn (src/SplittingMethod2D.cpp:124-141)."""
        f = FourierTransform((self.fNx, self.fNy), False)
        tmp = N.DeviceArray((self.fNx, self.fNy), np.complex128)
        f.transform(self.psi, tmp)
        Dkx = 2 * math.pi / (self.fDx * self.fNx)
        Dky = 2 * math.pi / (self.fDy * self.fNy)
        DTx = (Dkx * self._hbar) ** 2 / (2 * self._mass)
        DTy = (Dky * self._hbar) ** 2 / (2 * self._mass)
        kin = N.reduce1("qsim_hip_abs2_k2d", tmp.ptr, DTx, self.fNx, DTy, self.fNy)
        return kin / N.reduce1("qsim_hip_norm2", tmp.ptr, self._n)


def advance_batched(evolvers, nsteps):
    """Advance many same-shape evolvers together: one launch per sweep for all."""
    if not evolvers:
        return
    arr = (ctypes.c_void_p * len(evolvers))(*[e._job for e in evolvers])
    N.call("qsim_hip_split_steps_batched", arr, len(evolvers), nsteps, None)
    for e in evolvers:
        e.fStep += nsteps


# ------------------------------------------------------------- perturbation
_PRECOND = {"Vellekoop": 1, "Hao1": 2, "Hao2": 3, "BornIndentity": 0}


class _Perturbation:
    """QuPerturbation2D/3D (src/Perturbation2DImpl.cpp, src/Perturbation3DImpl.cpp)
    for one or many energies sharing the grid and V."""

    def _setup(self, dim, shape, lo, hi, V, energies, epsilon, direction, met, mass, hbar,
               opts):
        _device_type(opts)
        if met != SolverMethod.BornSerise:
            raise RuntimeError("not support method!")
        pre = bool(opts.Get("preconditional", False))
        name = opts.Get("preconditioner", "Vellekoop")
        if pre and name not in _PRECOND:
            raise RuntimeError("Unknown preconditioner")
        self._order = int(opts.Get("order", 1))
        self._dim = dim
        self._shape = tuple(shape)
        self._lo, self._hi = tuple(lo), tuple(hi)
        self._mass, self._hbar = float(mass), float(hbar)
        self._energies = [float(e) for e in np.atleast_1d(energies)]
        self._n = int(np.prod(shape))
        V = np.asarray(V, np.complex128).reshape(shape)
        if np.isnan(V.real).any() or np.isnan(V.imag).any():
            raise RuntimeError("nan in function")
        self._V = N.DeviceArray.from_host(V)
        d = N.BornDesc()
        d.dim = dim
        for a in range(3):
            d.n[a] = shape[a] if a < dim else 1
            d.lo[a] = lo[a] if a < dim else 0.0
            d.hi[a] = hi[a] if a < dim else 1.0
            d.direction[a] = direction[a] if a < dim else 0.0
        d.mass, d.hbar = self._mass, self._hbar
        d.preconditional = int(pre)
        d.preconditioner = _PRECOND.get(name, 1)
        d.slow = float(opts.Get("slow", 1.0))
        d.epsilon = float(epsilon)
        d.precision = _precision(opts)
        self._desc = d
        E = (ctypes.c_double * len(self._energies))(*self._energies)
        p = ctypes.c_void_p()
        N.call("qsim_hip_born_create", ctypes.byref(p), ctypes.byref(d), self._V.ptr, E,
               len(self._energies))
        self._job = p.value
        self._buf = N.DeviceArray(self._shape, np.complex128)
        self._last = None
        self._delta_norm = 0.0
        self.iters = [0] * len(self._energies)
        self.residual = [0.0] * len(self._energies)

    def Compute(self, max_iter=None, tol=0.0):
        """`order` iterations (reference semantics), or up to max_iter with an
        FP64 relative-update stopping rule when tol > 0."""
        k = len(self._energies)
        it = (ctypes.c_int * k)()
        rs = (ctypes.c_double * k)()
        if self._dim == 2:  # the Copy of fLastPsi, src/Perturbation2DImpl.cpp:56
            if self._last is None:
                self._last = N.DeviceArray(self._shape, np.complex128)
            N.call("qsim_hip_born_get_dpsi", self._job, 0, self._last.ptr, None)
        N.call("qsim_hip_born_iterate", self._job, self._order if max_iter is None else max_iter,
               float(tol), it, rs, None)
        self.iters = list(it)
        self.residual = list(rs)
        if self._dim == 2:
            # sqrt(MinusNorm2(psi, last) / Norm2(psi)), src/Perturbation2DImpl.cpp:80-82,
            # reduced on the device
            N.call("qsim_hip_born_get_dpsi", self._job, 0, self._buf.ptr, None)
            n = self._buf.size
            den = N.reduce1("qsim_hip_norm2", self._buf.ptr, n)
            num = N.reduce1("qsim_hip_minus_norm2", self._buf.ptr, self._last.ptr, n)
            self._delta_norm = math.sqrt(num / den) if den else 0.0
        return self

    def GetPsi(self, e=0):
        """scattered part dpsi of energy e"""
        N.call("qsim_hip_born_get_dpsi", self._job, e, self._buf.ptr, None)
        return self._buf.download()

    def GetVabsb(self, e=0):
        eps = ctypes.c_double()
        N.call("qsim_hip_born_get_veff", self._job, e, self._buf.ptr, ctypes.byref(eps), None)
        return self._buf.download()

    def GetEpsilon(self, e=0):
        eps = ctypes.c_double()
        N.call("qsim_hip_born_get_veff", self._job, e, self._buf.ptr, ctypes.byref(eps), None)
        return eps.value

    def GetEnergy(self, e=0):
        return self._energies[e]

    def GetMomentum(self, e=0):
        return math.sqrt(2 * self._mass * self._energies[e])

    def _xsec(self, e, c):
        ds = ctypes.c_double()
        amp = N.c128()
        N.call("qsim_hip_born_xsection", self._job, e, c[0], c[1], c[2], ctypes.byref(ds),
               ctypes.byref(amp), None)
        return ds.value, complex(amp.re, amp.im)

    def close(self):
        if getattr(self, "_job", None):
            N.lib().qsim_hip_born_destroy(self._job)
            self._job = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class QuPerturbation2D(_Perturbation):
    def init(self, v, x0, x1, nx, y0, y1, ny, en, epsilon, directionx, directiony, met, mass,
             hbar, opts):
        """QuPerturbation2D::init (include/QuSim.h:750-766)."""
        dx, dy = (x1 - x0) / nx, (y1 - y0) / ny
        X, Y = np.meshgrid(x0 + dx * np.arange(nx), y0 + dy * np.arange(ny), indexing="ij")
        V = v if isinstance(v, np.ndarray) else _eval2(v, X, Y)
        self._setup(2, (nx, ny), (x0, y0), (x1, y1), V, en, epsilon, (directionx, directiony),
                    met, mass, hbar, opts)
        return self

    def ComputeXSection(self, cosx, cosy, e=0):
        return self._xsec(e, (cosx, cosy, 0.0))[0]

    def ComputeTotalXSection(self, n, e=0):
        """src/ScatteringSolverImpl.cpp:199-212"""
        t = 0.0
        for i in range(n):
            th = i * 2 * math.pi / n
            t += self.ComputeXSection(math.cos(th), math.sin(th), e)
        return t * 2 * math.pi / n

    def GetDeltaPsiNorm(self):
        return self._delta_norm


class QuPerturbation3D(_Perturbation):
    def init(self, v, x0, x1, nx, y0, y1, ny, z0, z1, nz, en, epsilon, directionx, directiony,
             directionz, met, mass, hbar, opts):
        """QuPerturbation3D::init (include/QuSim.h:808-828)."""
        dx, dy, dz = (x1 - x0) / nx, (y1 - y0) / ny, (z1 - z0) / nz
        X, Y, Z = np.meshgrid(x0 + dx * np.arange(nx), y0 + dy * np.arange(ny),
                              z0 + dz * np.arange(nz), indexing="ij")
        V = v if isinstance(v, np.ndarray) else _eval3(v, X, Y, Z)
        self._setup(3, (nx, ny, nz), (x0, y0, z0), (x1, y1, z1), V, en, epsilon,
                    (directionx, directiony, directionz), met, mass, hbar, opts)
        return self

    def ComputeXSection(self, cosx, cosy, cosz, e=0):
        return self._xsec(e, (cosx, cosy, cosz))[0]

    def ScatteringAmplitude(self, cosx, cosy, cosz, e=0):
        """the raw far-field sum r of Xsection3D (src/DeviceTemplate.h:1646-1665)"""
        return self._xsec(e, (cosx, cosy, cosz))[1]

    def ComputeTotalXSection(self, npsi, ntheta, e=0):
        """src/ScatteringSolverImpl.cpp:363-379, theta indexed by j."""
        t = 0.0
        for j in range(ntheta):
            for i in range(npsi):
                phi = i * 2 * math.pi / npsi
                th = j * math.pi / ntheta
                t += self.ComputeXSection(math.sin(th) * math.cos(phi),
                                          math.sin(th) * math.sin(phi), math.cos(th), e) * math.sin(th)
        return t * 4 * math.pi / (ntheta * npsi)
