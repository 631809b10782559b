"""Caller-owned device buffers for the tests (hipMalloc / hipMemcpy through
libamdhip64, no torch): inputs and outputs that an entry point must accept in
device memory."""
import ctypes as C

import numpy as np

_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        _HIP = C.CDLL("libamdhip64.so")
        _HIP.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP.hipFree.argtypes = [C.c_void_p]
    return _HIP


class DevArray:
    """A caller-owned device copy of a float64 array (hipMalloc in the test)."""

    def __init__(self, a):
        self.host = np.asfortranarray(np.asarray(a, dtype=np.float64))
        self.ptr = C.c_void_p()
        assert _hip().hipMalloc(C.byref(self.ptr), max(self.host.nbytes, 8)) == 0
        assert _hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0  # host -> device

    def get(self, shape):
        out = np.empty(shape, dtype=np.float64, order="F")
        assert _hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device -> host
        return out

    def free(self):
        if self.ptr:
            _hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()
