"""Device buffers and streams for the GPU tests, through the HIP runtime
that oxdna_amd/liboxh.so itself links (libamdhip64.so.7, already loaded
once the library is): the tests play the caller that owns device memory,
as the reference's backend does for CUDABaseInteraction::compute_forces.
No second GPU runtime is brought into the process."""
from __future__ import annotations

import ctypes as C

import numpy as np

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        from oxdna_amd import _lib
        _lib.lib()   # loads liboxh.so and, with it, its HIP runtime
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipStreamDestroy.argtypes = [C.c_void_p]
        L.hipStreamSynchronize.argtypes = [C.c_void_p]
        L.hipDeviceSynchronize.argtypes = []
        L.hipSetDevice.argtypes = [C.c_int]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemset", "hipStreamCreate", "hipStreamDestroy",
                  "hipStreamSynchronize", "hipDeviceSynchronize", "hipSetDevice"):
            getattr(L, f).restype = C.c_int
        _HIP = L
    return _HIP


H2D, D2H = 1, 2


def _ok(status, what):
    if status != 0:
        raise RuntimeError(f"{what} failed with HIP status {status}")


class DeviceArray:
    """A device copy of a numpy array (same dtype and shape)."""

    def __init__(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        self.shape, self.dtype = a.shape, a.dtype
        self.nbytes = a.nbytes
        p = C.c_void_p()
        _ok(hip().hipSetDevice(0), "hipSetDevice")
        _ok(hip().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    @classmethod
    def zeros(cls, shape, dtype):
        return cls(np.zeros(shape, dtype))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if getattr(self, "ptr", None):
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A HIP stream of default priority."""

    def __init__(self):
        s = C.c_void_p()
        _ok(hip().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def close(self):
        if self.handle:
            hip().hipStreamDestroy(self.handle)
            self.handle = None


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
