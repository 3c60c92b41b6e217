"""ctypes binding of libqsim_hip.so (the C ABI declared in include/qsim_hip.h).

The HIP library is required: importing this module on a machine without the
built library, or calling into it without a GPU, raises -- there is no CPU
fallback anywhere in the product path."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqsim_hip.so")

_VP = ctypes.c_void_p
_SZ = ctypes.c_size_t
_D = ctypes.c_double
_I = ctypes.c_int


class QsimError(RuntimeError):
    pass


class c128(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


def C128(z):
    z = complex(z)
    return c128(z.real, z.imag)


class SplitDesc(ctypes.Structure):
    _fields_ = [("dim", _I), ("nx", _SZ), ("ny", _SZ), ("x0", _D), ("x1", _D), ("y0", _D),
                ("y1", _D), ("dt_re", _D), ("dt_im", _D), ("mass", _D), ("hbar", _D),
                ("method", _I), ("boundary", _I), ("precision", _I)]


class BornDesc(ctypes.Structure):
    _fields_ = [("dim", _I), ("n", _SZ * 3), ("lo", _D * 3), ("hi", _D * 3),
                ("direction", _D * 3), ("mass", _D), ("hbar", _D), ("preconditional", _I),
                ("preconditioner", _I), ("slow", _D), ("epsilon", _D), ("precision", _I)]


# name -> argtypes (all return int status unless listed in _RESTYPE)
_SIGS = {
    "qsim_hip_init": [_I],
    "qsim_hip_device_info": [ctypes.c_char_p, _SZ, ctypes.POINTER(_I), ctypes.POINTER(_SZ),
                             ctypes.POINTER(_SZ)],
    "qsim_hip_synchronize": [_VP],
    "qsim_hip_stream_create": [ctypes.POINTER(_VP)],
    "qsim_hip_stream_destroy": [_VP],
    "qsim_hip_malloc": [ctypes.POINTER(_VP), _SZ],
    "qsim_hip_free": [_VP],
    "qsim_hip_memcpy_h2d": [_VP, _VP, _SZ],
    "qsim_hip_memcpy_d2h": [_VP, _VP, _SZ],
    "qsim_hip_memcpy_d2d": [_VP, _VP, _SZ, _VP],
    "qsim_hip_event_create": [ctypes.POINTER(_VP)],
    "qsim_hip_event_destroy": [_VP],
    "qsim_hip_event_record": [_VP, _VP],
    "qsim_hip_event_elapsed_ms": [_VP, _VP, ctypes.POINTER(ctypes.c_float)],
    "qsim_hip_event_query": [_VP, ctypes.POINTER(_I)],
    "qsim_hip_event_synchronize": [_VP],
    "qsim_hip_set_zero": [_VP, _SZ, _VP],
    "qsim_hip_set_one": [_VP, _SZ, _VP],
    "qsim_hip_copy": [_VP, _VP, _SZ, _VP],
    "qsim_hip_copy_real": [_VP, _VP, _SZ, _VP],
    "qsim_hip_scale": [_VP, _D, _SZ, _VP],
    "qsim_hip_scale_c": [_VP, c128, _SZ, _VP],
    "qsim_hip_scale_to": [_VP, _VP, _D, _SZ, _VP],
    "qsim_hip_mul": [_VP, _VP, _SZ, _VP],
    "qsim_hip_mul3": [_VP, _VP, _VP, _SZ, _VP],
    "qsim_hip_exp_real": [_VP, _VP, c128, _SZ, _VP],
    "qsim_hip_exp_complex": [_VP, _VP, c128, _SZ, _VP],
    "qsim_hip_mul_exp_real": [_VP, c128, _VP, _SZ, _VP],
    "qsim_hip_mul_exp_complex": [_VP, c128, _VP, _SZ, _VP],
    "qsim_hip_mul_exp_k1d": [_VP, c128, _SZ, _VP],
    "qsim_hip_mul_exp_k2d": [_VP, c128, _D, _SZ, _D, _SZ, _VP],
    "qsim_hip_mul_exp_k3d": [_VP, c128, _D, _SZ, _D, _SZ, _D, _SZ, _VP],
    "qsim_hip_mul_exp_idx2": [_VP, c128, _SZ, _VP],
    "qsim_hip_copy_reverse_minus": [_VP, _VP, _SZ, _VP],
    "qsim_hip_set_elem": [_VP, _SZ, c128, _VP],
    "qsim_hip_get_elem": [_VP, _SZ, ctypes.POINTER(c128), _VP],
    "qsim_hip_exp_i_real": [_VP, _VP, _D, _SZ, _VP],
    "qsim_hip_sub": [_VP, _VP, _VP, _SZ, _VP],
    "qsim_hip_mul_r": [_VP, _VP, _VP, _SZ, _VP],
    "qsim_hip_mul_minus_i": [_VP, _VP, _VP, _SZ, _VP],
    "qsim_hip_mul_real_c": [_VP, _VP, _VP, _SZ, _VP],
    "qsim_hip_mul_real_alpha_c": [_VP, _VP, c128, _VP, _SZ, _VP],
    "qsim_hip_mul_exp3": [_VP, _VP, c128, _VP, _SZ, _VP],
    "qsim_hip_mul_exp_i_real": [_VP, _D, _VP, _SZ, _VP],
    "qsim_hip_scale_real": [_VP, _D, _SZ, _VP],
    "qsim_hip_scale_real_to": [_VP, _VP, _D, _SZ, _VP],
    "qsim_hip_scale_c_to": [_VP, _VP, c128, _SZ, _VP],
    "qsim_hip_set_elem_real": [_VP, _SZ, _D, _VP],
    "qsim_hip_get_elem_real": [_VP, _SZ, ctypes.POINTER(_D), _VP],
    "qsim_hip_abs2_idx_real": [_VP, _SZ, ctypes.POINTER(_D), _VP],
    "qsim_hip_g01d": [_VP, _D, _D, _D, _SZ, _VP],
    "qsim_hip_g02d": [_VP, _D, _D, _D, _D, _SZ, _SZ, _VP],
    "qsim_hip_g03d": [_VP, _D, _D, _D, _D, _D, _SZ, _SZ, _SZ, _VP],
    "qsim_hip_add_mul": [_VP, _VP, _VP, _SZ, _VP],
    "qsim_hip_add3": [_VP, _VP, _VP, _VP, _D, _SZ, _VP],
    "qsim_hip_linear_update": [_VP, _VP, _D, _SZ, _VP],
    "qsim_hip_vellekoop": [_VP, _VP, _VP, _D, _D, _SZ, _VP],
    "qsim_hip_hao1": [_VP, _VP, _VP, _D, _D, _SZ, _VP],
    "qsim_hip_hao2": [_VP, _VP, _VP, _D, _D, _SZ, _VP],
    "qsim_hip_norm2": [_VP, _SZ, ctypes.POINTER(_D), _VP],
    "qsim_hip_norm2_real": [_VP, _SZ, ctypes.POINTER(_D), _VP],
    "qsim_hip_minus_norm2": [_VP, _VP, _SZ, ctypes.POINTER(_D), _VP],
    "qsim_hip_max_abs": [_VP, _SZ, ctypes.POINTER(_D), _VP],
    "qsim_hip_min_abs": [_VP, _SZ, ctypes.POINTER(_D), _VP],
    "qsim_hip_sum_real": [_VP, _SZ, ctypes.POINTER(_D), _VP],
    "qsim_hip_sum_imag": [_VP, _SZ, ctypes.POINTER(_D), _VP],
    "qsim_hip_abs2_idx": [_VP, _SZ, ctypes.POINTER(_D), _VP],
    "qsim_hip_abs2_idx2": [_VP, _SZ, ctypes.POINTER(_D), _VP],
    "qsim_hip_abs2_k1d": [_VP, _D, _SZ, ctypes.POINTER(_D), _VP],
    "qsim_hip_abs2_k2d": [_VP, _D, _SZ, _D, _SZ, ctypes.POINTER(_D), _VP],
    "qsim_hip_abs2_k3d": [_VP, _D, _SZ, _D, _SZ, _D, _SZ, ctypes.POINTER(_D), _VP],
    "qsim_hip_abs2_mul_real": [_VP, _VP, _SZ, ctypes.POINTER(_D), _VP],
    "qsim_hip_abs2_mul_complex": [_VP, _VP, _SZ, ctypes.POINTER(c128), _VP],
    "qsim_hip_xsection3d": [_VP, _VP, _VP, _D, _D, _D, _D, _D, _D, _D, _D, _D, _SZ, _SZ, _SZ,
                            ctypes.POINTER(c128), _VP],
    "qsim_hip_xsection2d": [_VP, _VP, _VP, _D, _D, _D, _D, _D, _D, _SZ, _SZ,
                            ctypes.POINTER(c128), _VP],
    "qsim_hip_fft_plan_create": [ctypes.POINTER(_VP), _I, ctypes.POINTER(_SZ), _I, _I],
    "qsim_hip_fft_exec": [_VP, _VP, _VP, _VP],
    "qsim_hip_fft_plan_destroy": [_VP],
    "qsim_hip_split_create": [ctypes.POINTER(_VP), ctypes.POINTER(SplitDesc), _VP, _VP, _VP],
    "qsim_hip_split_steps": [_VP, _SZ, _VP],
    "qsim_hip_split_steps_timed": [_VP, _SZ, ctypes.POINTER(_D), ctypes.POINTER(_I), _VP],
    "qsim_hip_split_set_variants": [_VP, _I, _I],
    "qsim_hip_split_use_graph": [_VP, _I],
    "qsim_hip_split_autotune": [_VP, _I, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                ctypes.POINTER(_D), ctypes.POINTER(_D), _VP],
    "qsim_hip_split_load_psi": [_VP, _VP],
    "qsim_hip_split_norm2": [_VP, ctypes.POINTER(_D), _VP],
    "qsim_hip_split_destroy": [_VP],
    "qsim_hip_split_steps_batched": [ctypes.POINTER(_VP), _I, _SZ, _VP],
    "qsim_hip_split_steps_batched_noalloc": [ctypes.POINTER(_VP), _I, _SZ, _VP],
    "qsim_hip_split_set_xsplit": [_VP, _I],
    "qsim_hip_split_group_create": [ctypes.POINTER(_VP), _I, _I, _I],
    "qsim_hip_split_group_add": [_VP, _VP, _SZ],
    "qsim_hip_split_group_remove": [_VP, _VP],
    "qsim_hip_split_group_advance": [_VP, _I, _I, _VP],
    "qsim_hip_split_group_status": [_VP, _VP, ctypes.POINTER(_I),
                                    ctypes.POINTER(ctypes.c_longlong)],
    "qsim_hip_split_group_info": [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                  ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_longlong)],
    "qsim_hip_split_group_destroy": [_VP],
    "qsim_hip_split_get_schedule": [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                    ctypes.POINTER(_I)],
    "qsim_hip_born_create": [ctypes.POINTER(_VP), ctypes.POINTER(BornDesc), _VP,
                             ctypes.POINTER(_D), _I],
    "qsim_hip_born_iterate": [_VP, _I, _D, ctypes.POINTER(_I), ctypes.POINTER(_D), _VP],
    "qsim_hip_born_get_dpsi": [_VP, _I, _VP, _VP],
    "qsim_hip_born_get_veff": [_VP, _I, _VP, ctypes.POINTER(_D), _VP],
    "qsim_hip_born_xsection": [_VP, _I, _D, _D, _D, ctypes.POINTER(_D), ctypes.POINTER(c128),
                               _VP],
    "qsim_hip_born_destroy": [_VP],
    "qsim_hip_split_autotune_batched": [_VP, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                        ctypes.POINTER(_D), ctypes.POINTER(_D), _VP],
    "qsim_hip_split_batch_create": [_VP, _I],
    "qsim_hip_split_batch_steps": [_VP, _VP, _I, _SZ, _VP],
    "qsim_hip_split_batch_destroy": [_VP],
    "qsim_hip_gen_potential": [_I, _VP, _VP, _VP, _VP, _I, _D, _D, _VP, _VP, _VP, _VP],
    "qsim_hip_gen_packet": [_I, _VP, _VP, _VP, _VP, _VP, _D, _I, _VP, _VP],
}

_lib = None


def lib():
    """Load libqsim_hip.so; raises when the library was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QsimError("libqsim_hip.so not built (run `python -m qusim_amd.build`): %s"
                            % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.qsim_hip_last_error.restype = ctypes.c_char_p
        L.qsim_hip_last_error.argtypes = []
        for name, args in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:  # surfaces as AttributeError at the first call
                continue
            fn.argtypes = args
            fn.restype = _I
        _lib = L
    return _lib


def exported_symbols():
    return ["qsim_hip_last_error"] + list(_SIGS)


def check(rc):
    if rc != 0:
        msg = lib().qsim_hip_last_error().decode(errors="replace")
        raise QsimError("qsim_hip error %d: %s" % (rc, msg))


def call(name, *args):
    check(getattr(lib(), name)(*args))


_initialised = False


def init(device=0):
    global _initialised
    if not _initialised:
        call("qsim_hip_init", device)
        _initialised = True


def device_info():
    init()
    name = ctypes.create_string_buffer(256)
    cu = _I()
    free = _SZ()
    total = _SZ()
    call("qsim_hip_device_info", name, 256, ctypes.byref(cu), ctypes.byref(free),
         ctypes.byref(total))
    return {"name": name.value.decode(), "cu": cu.value, "mem_free": free.value,
            "mem_total": total.value}


def synchronize(stream=None):
    call("qsim_hip_synchronize", stream)


class DeviceArray:
    """Device allocation owned by Python (freed on close / garbage collection)."""

    def __init__(self, shape, dtype):
        init()
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if np.ndim(shape) else (
            int(shape),)
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        p = _VP()
        call("qsim_hip_malloc", ctypes.byref(p), self.nbytes)
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.size
        call("qsim_hip_memcpy_h2d", self.ptr, a.ctypes.data, self.nbytes)

    def download(self):
        out = np.empty(self.shape, self.dtype)
        call("qsim_hip_memcpy_d2h", out.ctypes.data, self.ptr, self.nbytes)
        return out

    def close(self):
        if getattr(self, "ptr", None):
            try:
                call("qsim_hip_free", self.ptr)
            finally:
                self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        init()
        p = _VP()
        call("qsim_hip_event_create", ctypes.byref(p))
        self.ptr = p.value

    def record(self, stream=None):
        call("qsim_hip_event_record", self.ptr, stream)

    def query(self):
        done = _I()
        call("qsim_hip_event_query", self.ptr, ctypes.byref(done))
        return bool(done.value)

    def synchronize(self):
        call("qsim_hip_event_synchronize", self.ptr)

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        call("qsim_hip_event_elapsed_ms", self.ptr, later.ptr, ctypes.byref(ms))
        return ms.value

    def __del__(self):
        try:
            if self.ptr:
                lib().qsim_hip_event_destroy(self.ptr)
        except Exception:
            pass


def reduce1(name, *args, stream=None):
    out = _D()
    call(name, *args, ctypes.byref(out), stream)
    return out.value


def reduce_c(name, *args, stream=None):
    out = c128()
    call(name, *args, ctypes.byref(out), stream)
    return complex(out.re, out.im)
