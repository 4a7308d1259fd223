"""Device buffers for GPU tests that call entry points taking device
pointers: hipMalloc'd through the HIP runtime libparallax_splat.so is linked
to (the test allocates; the entry points under test do not)."""
import ctypes as C

import numpy as np

from parallax_amd import splat as S

H2D, D2H = 1, 2


class DeviceMemory:
    def __init__(self):
        S.lib()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMalloc.restype = C.c_int
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipFree.restype = C.c_int
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipMemcpy.restype = C.c_int
        self.hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.hip.hipMemset.restype = C.c_int
        self.ptrs = []

    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.ptrs.append(p.value)
        return p.value

    def put(self, dev: int, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert self.hip.hipMemcpy(C.c_void_p(dev), a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def get(self, dev: int, a: np.ndarray):
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(dev), a.nbytes, D2H) == 0
        return a

    def poison(self, dev: int, nbytes: int):
        assert self.hip.hipMemset(C.c_void_p(dev), 0xFF, C.c_size_t(nbytes)) == 0  # NaN floats

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(C.c_void_p(p))
        self.ptrs = []
