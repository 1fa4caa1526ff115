"""Minimal device-memory helpers over the HIP runtime (ctypes), so callers of
the *_device entry points can keep inputs resident in HBM without pulling in
a framework."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                                                    "libamdhip64.so")):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("HIP runtime (libamdhip64.so) not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipSetDevice.argtypes = [C.c_int]
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class DeviceArray:
    """float64 device buffer of a fixed shape."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        self.ptr = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(self.ptr), max(8, self.nbytes)), "hipMalloc")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.shape)
        _ok(hip().hipMemcpy(d.ptr, a.ctypes.data_as(C.c_void_p), d.nbytes, 1), "hipMemcpy H2D")
        return d

    def upload(self, a):
        """Copy a host array of the same size into this buffer."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size * 8 != self.nbytes:
            raise ValueError("size mismatch")
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes, 1), "hipMemcpy H2D")

    def copy_from(self, other):
        """Device-to-device copy of a buffer of the same size."""
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _ok(hip().hipMemcpy(self.ptr, other.ptr, self.nbytes, 3), "hipMemcpy D2D")

    def to_host(self):
        out = np.empty(self.shape)
        _ok(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
