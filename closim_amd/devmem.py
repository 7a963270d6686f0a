"""Device buffers for the entry points that take device pointers, allocated,
copied and freed through the cloth library's own HIP runtime
(cs_device_alloc / cs_device_copy / cs_device_free), so they belong to the
runtime that runs the kernels even when another copy of the HIP runtime
(torch's) is loaded in the same process.  No CPU fallback: without a HIP
device every call raises."""
import ctypes as C

import numpy as np

from . import _abi

H2D, D2H, D2D = 1, 2, 3


def _lib():
    from .solver import load_library
    return load_library()


def _ok(rc):
    if rc != _abi.CS_OK:
        from .solver import ClothError
        raise ClothError(rc, _lib().cs_last_error().decode(errors="replace"))


class DeviceArray:
    """A device copy of a numpy array (or an uninitialised buffer of `shape`)."""

    def __init__(self, arr=None, dtype=None, shape=None, device=0):
        self.device = device
        if arr is not None:
            arr = np.ascontiguousarray(arr)
            self.dtype, self.shape = arr.dtype, arr.shape
        else:
            self.dtype, self.shape = np.dtype(dtype), tuple(shape)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        _ok(_lib().cs_device_alloc(device, max(self.nbytes, 1), C.byref(self.ptr)))
        if arr is not None and self.nbytes:
            _ok(_lib().cs_device_copy(device, self.ptr, C.c_void_p(arr.ctypes.data), self.nbytes, H2D))

    def numpy(self, count=None):
        """The buffer's contents, or its first `count` rows."""
        out = np.empty(self.shape, self.dtype)
        row = self.dtype.itemsize * int(np.prod(self.shape[1:], dtype=np.int64))
        nb = self.nbytes if count is None else min(self.nbytes, int(count) * row)
        if nb:
            _ok(_lib().cs_device_copy(self.device, C.c_void_p(out.ctypes.data), self.ptr, nb, D2H))
        return out if count is None else out[:count]

    def free(self):
        if self.ptr:
            _lib().cs_device_free(self.device, self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
