"""Device buffers for tests of the device-pointer entry points, through the
HIP runtime that libgmxcorr_hip.so itself loaded (libamdhip64.so.7).  torch
ships its own HIP runtime, which cannot share a process with the library's,
so the tests do not use torch tensors for device memory."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from gmx_correlation_amd import _lib
        _lib.load()                       # the library's runtime is loaded first
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


class DevArray:
    """A device copy of a numpy array (or an uninitialised one)."""

    def __init__(self, host=None, shape=None, dtype=None):
        if host is not None:
            host = np.ascontiguousarray(host)
            shape, dtype = host.shape, host.dtype
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed: {rc}")
        self.ptr = p.value
        if host is not None:
            self.upload(host)

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.nbytes == self.nbytes
        rc = hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, 1)   # host to device
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed: {rc}")

    def download(self):
        out = np.empty(self.shape, self.dtype)
        rc = hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2)   # device to host
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed: {rc}")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
