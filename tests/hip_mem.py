"""Device buffers for tests of entry points that take device pointers.

hipMalloc / hipMemcpy are called through the HIP runtime that libmhmsw.so
itself links, so the process holds one HIP runtime and the pointers are valid
for the library.  Test infrastructure only.
"""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from mhm2_staging_amd import _native
        _native.lib()  # loads libmhmsw.so and with it its HIP runtime
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
            f.restype = ctypes.c_int
        _hip = h
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """A device copy of a numpy array (or an uninitialised one of nbytes)."""

    def __init__(self, host=None, nbytes=None):
        h = hip()
        self.nbytes = int(host.nbytes if host is not None else nbytes)
        self.ptr = ctypes.c_void_p()
        rc = h.hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1))
        if rc != 0:
            raise RuntimeError(f"hipMalloc({self.nbytes}) failed: {rc}")
        if host is not None and self.nbytes:
            host = np.ascontiguousarray(host)
            rc = h.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D)
            if rc != 0:
                raise RuntimeError(f"hipMemcpy H2D failed: {rc}")

    def to_host(self, dtype, count):
        out = np.zeros(count, dtype=dtype)
        if out.nbytes:
            rc = hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H)
            if rc != 0:
                raise RuntimeError(f"hipMemcpy D2H failed: {rc}")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def synchronize():
    rc = hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")
