"""Device buffers for GPU tests through the HIP runtime that libinfraga_gpu.so
itself links (libamdhip64.so.7 of /opt/rocm), called with ctypes.

torch bundles a second HIP/HSA runtime; a test that initialises it after the
engine's runtime has been working in the same process can find no device, so
tests that need raw device pointers take them from here instead.
"""
import ctypes as C

import numpy as np

H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        from infraga_amd import engine
        engine.load_library()                  # loads the runtime the engine uses
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.L.hipDeviceSynchronize.argtypes = []
        self.L.hipSetDevice.argtypes = [C.c_int]
        self.live = []

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def alloc(self, nbytes, fill=None):
        p = C.c_void_p()
        self._ok(self.L.hipMalloc(C.byref(p), max(int(nbytes), 1)), "hipMalloc")
        self.live.append(p.value)
        if fill is not None:
            self._ok(self.L.hipMemset(p, fill, max(int(nbytes), 1)), "hipMemset")
        return p.value

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        p = self.alloc(a.nbytes)
        self._ok(self.L.hipMemcpy(C.c_void_p(p), a.ctypes.data_as(C.c_void_p), a.nbytes, H2D), "hipMemcpy H2D")
        return p

    def download(self, ptr, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            self._ok(self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def sync(self):
        self._ok(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free_all(self):
        self.sync()
        for p in self.live:
            self.L.hipFree(C.c_void_p(p))
        self.live = []
