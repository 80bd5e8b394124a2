"""Device buffers for tests that call the C ABI with device pointers:
hipMalloc / hipMemcpy / hipFree of the HIP runtime liboagpu.so itself links
(/opt/rocm/lib/libamdhip64.so), through ctypes.  (torch's wheel bundles a
HIP runtime of its own, which cannot share the GPU with the library's in one
process.)"""
import ctypes as C
import os

import numpy as np

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        import openaccel_amd as oa
        oa.load_library()                       # the runtime liboagpu.so loaded
        path = "libamdhip64.so.7"
        for cand in ("/opt/rocm/lib/libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
            if os.path.exists(cand):
                path = cand
                break
        L = C.CDLL(path)
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        for f in (L.hipMalloc, L.hipFree, L.hipMemcpy, L.hipDeviceSynchronize):
            f.restype = C.c_int
        _hip = L
    return _hip


class DeviceArray:
    """A device copy of a float64 array (hipMalloc'd; freed on close)."""

    def __init__(self, shape, host=None, fill=np.nan):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 8)) == 0
        src = np.full(self.shape, fill) if host is None else np.ascontiguousarray(host, dtype=np.float64)
        assert src.shape == self.shape
        assert hip().hipMemcpy(self.ptr, src.ctypes.data, self.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.shape)
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def close(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        self.close()
