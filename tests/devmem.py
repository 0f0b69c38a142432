"""Device buffers for GPU tests, allocated through the HIP runtime that
libbnzgpu.so itself is linked against (found in this process's mappings
after the library is loaded), so that tests need no second runtime."""
import ctypes

import numpy as np

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        from bonzo_amd._lib import load
        load()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line:
                    path = line.split()[-1]
                    break
        if path is None:
            raise RuntimeError("libamdhip64 is not mapped after loading libbnzgpu.so")
        L = ctypes.CDLL(path)
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        for name in ("hipMalloc", "hipFree", "hipMemcpy"):
            getattr(L, name).restype = ctypes.c_int
        _hip = L
    return _hip


class DeviceArray:
    """A float64 device copy of a host array; ``get()`` copies it back."""

    def __init__(self, host):
        self.host = np.ascontiguousarray(host, dtype=np.float64)
        self.ptr = ctypes.c_void_p()
        H = hip()
        assert H.hipMalloc(ctypes.byref(self.ptr), max(8, self.host.nbytes)) == 0
        assert H.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr.value:
            hip().hipFree(self.ptr)
        self.ptr = ctypes.c_void_p()
