"""Device buffers for tests that call the C ABI's device-pointer entry points (the prefix scans).

Allocated through the HIP runtime libdeme.so itself links (loaded first, so the process holds one runtime), not
through torch: torch's wheel brings its own HIP runtime of another release under the same soname.
"""
import ctypes as C

import numpy as np

from deme_amd.abi import load

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        load()  # libdeme.so first: its runtime is the one the process uses
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipMalloc.restype = C.c_int
        h.hipFree.argtypes = [C.c_void_p]
        h.hipFree.restype = C.c_int
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpy.restype = C.c_int
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipMemset.restype = C.c_int
        h.hipDeviceSynchronize.restype = C.c_int
        _hip = h
    return _hip


class DeviceArray:
    """A device allocation of `nbytes` (at least 1), filled with byte `fill`; freed by close()."""

    def __init__(self, nbytes, fill=0xFF):
        self.nbytes = max(int(nbytes), 1)
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), self.nbytes) == 0, "hipMalloc failed"
        assert hip().hipMemset(self.ptr, fill, self.nbytes) == 0
        assert hip().hipDeviceSynchronize() == 0

    @classmethod
    def of(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if a.nbytes:
            assert hip().hipMemcpy(d.ptr, C.c_void_p(a.ctypes.data), a.nbytes, H2D) == 0
        return d

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        if out.nbytes:
            assert hip().hipDeviceSynchronize() == 0
            assert hip().hipMemcpy(C.c_void_p(out.ctypes.data), self.ptr, out.nbytes, D2H) == 0
        return out

    def close(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()
