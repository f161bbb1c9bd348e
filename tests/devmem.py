"""Device buffers for tests that call C-ABI entry points taking device
pointers (HIP runtime through ctypes; no torch needed)."""
import ctypes as C
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise OSError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """n doubles (or the contents of a host array) in device memory."""

    def __init__(self, host=None, n=None):
        if host is not None:
            host = np.ascontiguousarray(host, dtype=np.float64)
            n = host.size
        self.n = int(n)
        self.ptr = C.c_void_p()
        rc = hip().hipMalloc(C.byref(self.ptr), max(self.n, 1) * 8)
        if rc != 0:
            raise RuntimeError("hipMalloc failed: %d" % rc)
        if host is not None:
            self._check(hip().hipMemcpy(self.ptr, host.ctypes.data, self.n * 8, H2D))

    @staticmethod
    def _check(rc):
        if rc != 0:
            raise RuntimeError("HIP call failed: %d" % rc)

    def get(self):
        out = np.zeros(self.n)
        self._check(hip().hipMemcpy(out.ctypes.data, self.ptr, self.n * 8, D2H))
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
