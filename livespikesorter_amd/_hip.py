"""Minimal device-memory plumbing for callers that hand the matcher
HBM-resident input (tests, bench).

It binds the HIP runtime instance that ``liblss.so`` itself is linked
against (``libamdhip64.so.7``, resolved by soname to the copy already mapped
into the process), so buffers allocated here live in the same runtime as the
library's own.  Loading a second, different HIP runtime into one process
(for example a framework's bundled copy after this one) leaves the later one
without devices; keep one runtime per process.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from ._capi import lib

H2D = 1
D2H = 2
D2D = 3

_H = None


def hip() -> ct.CDLL:
    global _H
    if _H is None:
        lib()                                    # maps liblss.so and its HIP runtime first
        h = ct.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
        h.hipFree.argtypes = [ct.c_void_p]
        h.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
        h.hipDeviceSynchronize.argtypes = []
        h.hipSetDevice.argtypes = [ct.c_int]
        h.hipGetErrorString.restype = ct.c_char_p
        h.hipGetErrorString.argtypes = [ct.c_int]
        _H = h
    return _H


def _check(code: int, what: str) -> None:
    if code != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(code).decode()}")


def synchronize() -> None:
    _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceBuffer:
    """A plain device allocation (hipMalloc) holding a copy of a numpy array."""

    def __init__(self, nbytes: int, device: int = 0):
        _check(hip().hipSetDevice(device), "hipSetDevice")
        self._p = ct.c_void_p()
        _check(hip().hipMalloc(ct.byref(self._p), max(1, int(nbytes))), "hipMalloc")
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a: np.ndarray, device: int = 0) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes, device)
        b.upload(a)
        return b

    @property
    def ptr(self) -> int:
        return int(self._p.value or 0)

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        if offset + a.nbytes > self.nbytes:
            raise ValueError("upload beyond the buffer")
        _check(hip().hipMemcpy(ct.c_void_p(self.ptr + offset), a.ctypes.data, a.nbytes, H2D), "hipMemcpy")

    def download(self, shape, dtype, offset: int = 0) -> np.ndarray:
        out = np.empty(shape, dtype)
        if offset + out.nbytes > self.nbytes:
            raise ValueError("download beyond the buffer")
        _check(hip().hipMemcpy(out.ctypes.data, ct.c_void_p(self.ptr + offset), out.nbytes, D2H), "hipMemcpy")
        return out

    def copy_from(self, other: "DeviceBuffer", nbytes: int) -> None:
        """Device-to-device copy (the runtime's blit kernel)."""
        if nbytes > self.nbytes or nbytes > other.nbytes:
            raise ValueError("copy beyond a buffer")
        _check(hip().hipMemcpy(ct.c_void_p(self.ptr), ct.c_void_p(other.ptr), nbytes, D2D), "hipMemcpy")

    def free(self) -> None:
        if self._p:
            hip().hipFree(self._p)
            self._p = ct.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
