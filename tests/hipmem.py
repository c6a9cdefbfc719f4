"""Device buffers for tests of the raw-pointer entry points: hipMalloc / hipMemcpy through
ctypes (libamdhip64), no torch."""
import ctypes as C

import numpy as np

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        h = C.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        _HIP = h
    return _HIP


class DeviceArray:
    """A device copy of a numpy array (or an uninitialised one of a shape and dtype)."""

    def __init__(self, host=None, shape=None, dtype=None, fill=None):
        if host is not None:
            host = np.ascontiguousarray(host)
            shape, dtype = host.shape, host.dtype
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 4)) == 0
        if host is not None and self.nbytes:
            assert hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, 1) == 0   # H2D
        elif fill is not None:
            assert hip().hipMemset(self.ptr, fill, max(self.nbytes, 4)) == 0

    def get(self):
        out = np.empty(self.shape, self.dtype)
        assert hip().hipDeviceSynchronize() == 0
        if self.nbytes:
            assert hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0   # D2H
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
