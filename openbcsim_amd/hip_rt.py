"""Minimal HIP runtime access through ctypes (device buffers for the
device-resident entry points).  Uses the same libamdhip64 that libbcsim_hip.so
links, so pointers are valid across both."""
from __future__ import annotations

import ctypes as C
import os

_hip = None


def hip():
    global _hip
    if _hip is None:
        for path in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
            try:
                _hip = C.CDLL(path)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetErrorString.restype = C.c_char_p
        _hip.hipGetErrorString.argtypes = [C.c_int]
    return _hip


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


class DeviceBuffer:
    """hipMalloc'd bytes, freed on close()/del."""

    def __init__(self, nbytes: int):
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        _check(hip().hipMalloc(C.byref(self.ptr), max(1, self.nbytes)), "hipMalloc")

    @property
    def address(self) -> int:
        return int(self.ptr.value)

    def to_host(self, out) -> None:
        """Copy into a writable numpy array (device-to-host, synchronous)."""
        _check(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, 2), "hipMemcpy D2H")

    def close(self):
        if self.ptr and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
