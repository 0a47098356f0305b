"""Minimal HIP runtime access through ctypes: device buffers and copies for
callers that keep the build's inputs and outputs resident in HBM."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        L = ctypes.CDLL("libamdhip64.so")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipSetDevice.argtypes = [ctypes.c_int]
        L.hipGetErrorString.restype = ctypes.c_char_p
        L.hipGetErrorString.argtypes = [ctypes.c_int]
        _hip = L
    return _hip


def _check(err, what):
    if err != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(err).decode()}")


class DeviceArray:
    """A float64 device buffer of `n` elements."""

    def __init__(self, n):
        self.n = int(n)
        p = ctypes.c_void_p()
        _check(hip().hipMalloc(ctypes.byref(p), max(8 * self.n, 8)), "hipMalloc")
        self.ptr = p.value

    def upload(self, a):
        """Copy a host array in; 2-D arrays are taken column-major."""
        a = np.asarray(a, dtype=np.float64)
        a = np.ascontiguousarray(a.ravel(order="F" if a.ndim == 2 else "C"))
        assert a.size == self.n
        _check(hip().hipMemcpy(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p),
                               8 * self.n, 1), "hipMemcpy H2D")

    def download(self, shape=None):
        """Copy out; a 2-D shape is filled column-major."""
        out = np.empty(self.n)
        _check(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                               8 * self.n, 2), "hipMemcpy D2H")
        return out.reshape(shape, order="F") if shape is not None else out

    def free(self):
        if self.ptr:
            hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy_device(dst, src, nbytes):
    """Device-to-device copy (hipMemcpyDeviceToDevice)."""
    _check(hip().hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, 3),
           "hipMemcpy D2D")


def synchronize():
    _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def memory_info():
    """(free, total) device bytes of the current device."""
    free = ctypes.c_size_t()
    total = ctypes.c_size_t()
    _check(hip().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value, total.value
