"""FourierTransform{1,2,3}D on the HIP device (reference src/FourierTransform.h:12-36):
unnormalised, forward sign -, inverse sign +, out of place."""
import ctypes

import numpy as np

from . import _native as N


class FourierTransform:
    """A plan over `dims` (outer -> inner).  precision 64 (complex128) or 32
    (complex64).  `transform(src, dst)` takes DeviceArray or raw device pointers."""

    def __init__(self, dims, inverse=False, precision=64):
        N.init()
        self.dims = tuple(int(d) for d in dims)
        self.inverse = bool(inverse)
        self.precision = int(precision)
        arr = (ctypes.c_size_t * len(self.dims))(*self.dims)
        p = ctypes.c_void_p()
        N.call("qsim_hip_fft_plan_create", ctypes.byref(p), len(self.dims), arr,
               int(self.inverse), self.precision)
        self.ptr = p.value

    @staticmethod
    def Create(dims, inverse, precision=64):
        return FourierTransform(dims, inverse, precision)

    def transform(self, src, dst, stream=None):
        s = src.ptr if hasattr(src, "ptr") else src
        d = dst.ptr if hasattr(dst, "ptr") else dst
        N.call("qsim_hip_fft_exec", self.ptr, s, d, stream)

    Transform = transform

    def __call__(self, host_array):
        """Convenience: host array in, host array out (synchronous)."""
        dt = np.complex128 if self.precision == 64 else np.complex64
        a = np.ascontiguousarray(host_array, dtype=dt).reshape(self.dims)
        src = N.DeviceArray.from_host(a)
        dst = N.DeviceArray(a.shape, dt)
        self.transform(src, dst)
        out = dst.download()
        src.close()
        dst.close()
        return out

    def __del__(self):
        try:
            if self.ptr:
                N.lib().qsim_hip_fft_plan_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


def fftn(a, inverse=False, precision=64):
    a = np.asarray(a)
    return FourierTransform(a.shape, inverse, precision)(a)
