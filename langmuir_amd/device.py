"""Device buffers from the HIP runtime that ``liblangmuir_gpu.so`` links.

The ``*_dev`` entry points of the C-ABI take device pointers.  A second HIP
runtime in the same process (torch's bundled one, for instance) would not share
the library's device state, so buffers come from ``libamdhip64`` directly.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None

H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceBuffer:
    """nbytes of device memory (rounded up to whole 4-byte words)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self._ptr = C.c_void_p()
        rc = hip().hipMalloc(C.byref(self._ptr), C.c_size_t(max((self.nbytes + 3) // 4 * 4, 4)))
        if rc != 0:
            raise MemoryError(f"hipMalloc({self.nbytes}) failed with {rc}")

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    @property
    def ptr(self) -> int:
        return int(self._ptr.value or 0)

    def at(self, offset_bytes: int) -> int:
        return self.ptr + int(offset_bytes)

    def upload(self, a: np.ndarray, offset_bytes: int = 0) -> None:
        a = np.ascontiguousarray(a)
        if offset_bytes + a.nbytes > self.nbytes:
            raise ValueError("upload past the end of the buffer")
        rc = hip().hipMemcpy(C.c_void_p(self.at(offset_bytes)), a.ctypes.data_as(C.c_void_p),
                             C.c_size_t(a.nbytes), H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed with {rc}")

    def download(self, dtype, count: int, offset_bytes: int = 0) -> np.ndarray:
        out = np.empty(int(count), dtype)
        if offset_bytes + out.nbytes > self.nbytes:
            raise ValueError("download past the end of the buffer")
        rc = hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.at(offset_bytes)),
                             C.c_size_t(out.nbytes), D2H)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed with {rc}")
        return out

    def free(self) -> None:
        if self._ptr.value:
            hip().hipFree(self._ptr)
            self._ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
