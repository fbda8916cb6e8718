"""Device buffers through the library's own C ABI (no torch)."""
import ctypes

import numpy as np


class Dev:
    def __init__(self, pkg, arr=None, nbytes=None):
        self.L = pkg.lib()
        self.L.rsvdt_hip_malloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.rsvdt_hip_memcpy_h2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        self.L.rsvdt_hip_memcpy_d2h.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        self.L.rsvdt_hip_free.argtypes = [ctypes.c_void_p]
        self.p = ctypes.c_void_p()
        self.nbytes = int(arr.nbytes if arr is not None else nbytes)
        assert self.L.rsvdt_hip_malloc(ctypes.byref(self.p), max(self.nbytes, 16)) == 0
        if arr is not None:
            a = np.ascontiguousarray(arr)
            assert self.L.rsvdt_hip_memcpy_h2d(self.p, a.ctypes.data_as(ctypes.c_void_p), self.nbytes) == 0

    def get(self, dtype, shape):
        out = np.zeros(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        assert self.L.rsvdt_hip_sync() == 0
        assert self.L.rsvdt_hip_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), self.p, out.nbytes) == 0
        return out

    def __del__(self):
        try:
            self.L.rsvdt_hip_free(self.p)
        except Exception:
            pass


def init(pkg):
    L = pkg.lib()
    rc = L.rsvdt_hip_init(0)
    assert rc == 0, L.rsvdt_hip_last_error()
    return L
