"""Device buffers for the GPU tests through the HIP runtime (ctypes on libamdhip64.so): the
torch build in this image does not see the GPU of the test machines, so device memory for the
device-pointer entry points comes from hipMalloc directly (test infrastructure)."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """A device copy of a numpy array (same dtype and shape)."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.shape, self.dtype, self.nbytes = host.shape, host.dtype, host.nbytes
        p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 16)) == 0
        self.ptr = p.value
        self.upload(host)

    @classmethod
    def full(cls, shape, dtype, value):
        return cls(np.full(shape, value, dtype))

    def upload(self, host):
        host = np.ascontiguousarray(host, self.dtype)
        assert host.nbytes == self.nbytes
        assert hip().hipMemcpy(ctypes.c_void_p(self.ptr), host.ctypes.data_as(ctypes.c_void_p), self.nbytes, H2D) == 0

    def download(self):
        out = np.empty(self.shape, self.dtype)
        assert hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr), self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    assert hip().hipDeviceSynchronize() == 0
