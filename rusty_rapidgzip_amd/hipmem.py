"""Device buffers through the same HIP runtime librrg_gpu.so links (/opt/rocm).

PyTorch's ROCm wheel bundles its own HIP/HSA runtime under another SONAME; two
runtimes in one process do not share the device, so the decoder's users move
bytes with this helper instead of torch tensors.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_hip = None
H2D, D2H, D2D = 1, 2, 3


def hip() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        from . import _native

        _native.gpu_lib()  # load the decoder first so both resolve to one runtime
        path = "/opt/rocm/lib/libamdhip64.so.7"
        _hip = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipSetDevice.argtypes = [ctypes.c_int]
    return _hip


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


def synchronize() -> None:
    _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceBuffer:
    """A hipMalloc'd byte buffer on the current device."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(hip().hipMalloc(ctypes.byref(p), max(1, self.nbytes)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, data) -> "DeviceBuffer":
        a = np.ascontiguousarray(np.frombuffer(data, dtype=np.uint8) if isinstance(data, (bytes, bytearray))
                                 else np.asarray(data, dtype=np.uint8))
        b = cls(a.size)
        if a.size:
            _check(hip().hipMemcpy(ctypes.c_void_p(b.ptr), a.ctypes.data, a.size, H2D), "hipMemcpy H2D")
        return b

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):  # pragma: no cover
        try:
            self.free()
        except Exception:
            pass


def to_host(ptr: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    if n:
        _check(hip().hipMemcpy(out.ctypes.data, ctypes.c_void_p(ptr), n, D2H), "hipMemcpy D2H")
    return out
