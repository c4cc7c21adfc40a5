"""Minimal ctypes access to the HIP runtime for device buffers and
synchronization (plumbing for the benchmark and the device-pointer tests)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_hip = None


def lib():
    global _hip
    if _hip is None:
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        for name in (os.path.join(rocm, "lib", "libamdhip64.so"), "libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        _hip.hipGetErrorString.restype = C.c_char_p
    return _hip


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {lib().hipGetErrorString(rc).decode()}")


def device_count() -> int:
    n = C.c_int(0)
    rc = lib().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def synchronize() -> None:
    _check(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A float64 buffer in HBM."""

    def __init__(self, count: int):
        self.count = int(count)
        self.ptr = C.c_void_p()
        _check(lib().hipMalloc(C.byref(self.ptr), max(self.count, 1) * 8), "hipMalloc")

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        d = cls(a.size)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        assert a.size == self.count
        _check(lib().hipMemcpy(self.ptr, a.ctypes.data, a.size * 8, 1), "hipMemcpy H2D")

    def copy_from(self, other: "DeviceArray") -> None:
        assert other.count == self.count
        _check(lib().hipMemcpy(self.ptr, other.ptr, self.count * 8, 3), "hipMemcpy D2D")

    def download(self) -> np.ndarray:
        out = np.empty(self.count)
        _check(lib().hipMemcpy(out.ctypes.data, self.ptr, self.count * 8, 2), "hipMemcpy D2H")
        return out

    def free(self) -> None:
        if self.ptr:
            lib().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
