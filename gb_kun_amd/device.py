"""Device buffers, streams and events for the Python host side.

They come from the HIP runtime that libgbk.so itself links (the soname it
was built against), so every pointer handed to the C ABI belongs to the same
runtime instance as the library's kernels and stream.  Timing events are
recorded on the library stream, where the kernels run.
"""
import ctypes as C
import os

import numpy as np

from . import lib as _gbk  # noqa: F401  (loads libgbk.so and with it the HIP runtime)


def _load_hip():
    for name in ("libamdhip64.so.7", "libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return C.CDLL(name, mode=getattr(os, "RTLD_NOLOAD", 0) | C.RTLD_GLOBAL)
        except OSError:
            continue
    for name in ("libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return C.CDLL(name, mode=C.RTLD_GLOBAL)
        except OSError:
            continue
    raise ImportError("the HIP runtime linked by libgbk.so could not be loaded")


hip = _load_hip()
_vp = C.c_void_p
for _name, _args in (("hipMalloc", [C.POINTER(_vp), C.c_size_t]), ("hipFree", [_vp]),
                     ("hipMemcpy", [_vp, _vp, C.c_size_t, C.c_int]),
                     ("hipMemset", [_vp, C.c_int, C.c_size_t]),
                     ("hipDeviceSynchronize", []), ("hipEventCreate", [C.POINTER(_vp)]),
                     ("hipEventDestroy", [_vp]), ("hipEventRecord", [_vp, _vp]),
                     ("hipEventSynchronize", [_vp]),
                     ("hipEventElapsedTime", [C.POINTER(C.c_float), _vp, _vp]),
                     ("hipMemcpyAsync", [_vp, _vp, C.c_size_t, C.c_int, _vp])):
    getattr(hip, _name).restype = C.c_int
    getattr(hip, _name).argtypes = _args

H2D, D2H, D2D = 1, 2, 3


def _ck(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


def synchronize():
    _ck(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A 1-D or 2-D device array of a numpy dtype (row-major)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._p = _vp()
        _ck(hip.hipMalloc(C.byref(self._p), max(self.nbytes, 256)), "hipMalloc")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            _ck(hip.hipMemcpy(d._p, a.ctypes.data_as(_vp), d.nbytes, H2D), "hipMemcpy H2D")
        return d

    @classmethod
    def zeros(cls, shape, dtype):
        d = cls(shape, dtype)
        _ck(hip.hipMemset(d._p, 0, max(d.nbytes, 256)), "hipMemset")
        return d

    @property
    def ptr(self):
        return self._p.value

    def numpy(self):
        """Synchronizes the device and copies the array back."""
        synchronize()
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            _ck(hip.hipMemcpy(out.ctypes.data_as(_vp), self._p, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self._p:
            hip.hipFree(self._p)
            self._p = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy_d2d(dst, src, nbytes=None):
    """Device-to-device copy on the library stream."""
    nb = min(dst.nbytes, src.nbytes) if nbytes is None else nbytes
    _ck(hip.hipMemcpyAsync(dst._p, src._p, nb, D2D, _vp(_gbk.gbk_get_stream())), "hipMemcpyAsync D2D")


class Event:
    def __init__(self):
        self._e = _vp()
        _ck(hip.hipEventCreate(C.byref(self._e)), "hipEventCreate")

    def record(self, stream=None):
        s = stream if stream is not None else _gbk.gbk_get_stream()
        _ck(hip.hipEventRecord(self._e, _vp(s)), "hipEventRecord")

    def synchronize(self):
        _ck(hip.hipEventSynchronize(self._e), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = C.c_float(0)
        _ck(hip.hipEventElapsedTime(C.byref(ms), self._e, end._e), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if self._e:
                hip.hipEventDestroy(self._e)
        except Exception:
            pass
