"""Device buffers for GPU tests that call entry points taking device pointers:
hipMalloc / hipMemcpy / hipFree of the HIP runtime the engine library itself
links (libamdhip64.so.7 from ROCm), called through ctypes.  Test plumbing only.
(torch ships a HIP runtime of its own; two runtimes in one process do not share
the device, so torch is not used for this.)"""
import ctypes as C

import numpy as np

_H2D, _D2H = 1, 2
_hip = None


def _lib():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: hipError {rc}")


class DeviceArray:
    """A device allocation with a numpy dtype and shape."""

    def __init__(self, shape, dtype, fill_byte=None):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _ok(_lib().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value
        if fill_byte is not None:
            _ok(_lib().hipMemset(self.ptr, fill_byte, self.nbytes), "hipMemset")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _ok(_lib().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, _H2D), "hipMemcpy H2D")
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _ok(_lib().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            _lib().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


def synchronize():
    _ok(_lib().hipDeviceSynchronize(), "hipDeviceSynchronize")
