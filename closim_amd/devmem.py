"""Device buffers through the HIP runtime (hipMalloc / hipMemcpy / hipFree)
for the entry points that take device pointers.  No CPU fallback: without a
HIP device every call raises."""
import ctypes as C

import numpy as np

_hip = None
_H2D, _D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        L = C.CDLL("libamdhip64.so")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipSetDevice.argtypes = [C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipGetErrorString.restype = C.c_char_p
        _hip = L
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


class DeviceArray:
    """A device copy of a numpy array (or an uninitialised buffer of `shape`)."""

    def __init__(self, arr=None, dtype=None, shape=None, device=0):
        _ok(hip().hipSetDevice(device), "hipSetDevice")
        if arr is not None:
            arr = np.ascontiguousarray(arr)
            self.dtype, self.shape = arr.dtype, arr.shape
        else:
            self.dtype, self.shape = np.dtype(dtype), tuple(shape)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")
        if arr is not None and self.nbytes:
            _ok(hip().hipMemcpy(self.ptr, arr.ctypes.data, self.nbytes, _H2D), "hipMemcpy H2D")

    def numpy(self, count=None):
        out = np.empty(self.shape, self.dtype)
        nb = self.nbytes if count is None else min(self.nbytes, int(count) * self.dtype.itemsize * int(
            np.prod(self.shape[1:], dtype=np.int64)))
        if nb:
            _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, nb, _D2H), "hipMemcpy D2H")
        return out if count is None else out[:count]

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
