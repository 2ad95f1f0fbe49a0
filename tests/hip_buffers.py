"""Device buffers for tests of the device-pointer entry points, through the
HIP runtime that libpmcd.so itself links (libamdhip64.so.7). Test plumbing
only: torch bundles another HIP runtime, and two runtimes in one process do
not share a device."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        import parmoon2022_amd as pm
        pm.lib()  # loads the runtime libpmcd.so links
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


class DeviceArray:
    """A device copy of a numpy array (freed on close or deletion)."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a)
        self.nbytes = max(self.a.nbytes, 8)
        p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(p), self.nbytes) == 0, "hipMalloc failed"
        self.ptr = p.value
        if self.a.nbytes:
            assert hip().hipMemcpy(self.ptr, self.a.ctypes.data, self.a.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.a)
        if out.nbytes:
            assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()
