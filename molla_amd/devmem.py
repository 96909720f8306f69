"""Device buffers through the HIP runtime (ctypes), for the entry points that
take device pointers: mpm_read_field_device, mpm_diff_backward_device and
mpm_sort_keys_noalloc.  The buffers belong to the caller, as in the C ABI;
the library never frees or resizes them.

Copies run on the null stream and wait for completion; a caller that filled
a buffer through a handle waits for that handle first (``synchronize``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None
_H2D, _D2H = 1, 2


def _rt():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
    return _hip


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class DeviceArray:
    """`shape` elements of `dtype` in device memory of the current device."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._p = C.c_void_p()
        _check(_rt().hipMalloc(C.byref(self._p), C.c_size_t(max(self.nbytes, 1))), "hipMalloc")

    @classmethod
    def from_host(cls, a) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    @property
    def address(self) -> int:
        """The device address, as the library's *_dev arguments take it."""
        return self._p.value

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _check(_rt().hipMemcpy(self._p, a.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), _H2D), "hipMemcpy")
        _check(_rt().hipDeviceSynchronize(), "hipDeviceSynchronize")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _check(_rt().hipMemcpy(out.ctypes.data_as(C.c_void_p), self._p, C.c_size_t(self.nbytes), _D2H),
               "hipMemcpy")
        return out

    def free(self):
        if self._p and self._p.value:
            _rt().hipFree(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()
