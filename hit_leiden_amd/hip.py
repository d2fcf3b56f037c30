"""Minimal ctypes binding of the HIP runtime (device buffers for batches that
must already sit in HBM when a timed region starts)."""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None
H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise OSError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipSetDevice.argtypes = [C.c_int]
        _hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """A host numpy array copied into device memory."""

    def __init__(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        self.nbytes = max(a.nbytes, 1)
        self.dtype = a.dtype
        self.size = a.size
        self.ptr = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(self.ptr), self.nbytes), "hipMalloc")
        if a.nbytes:
            _ok(hip().hipMemcpy(self.ptr, C.c_void_p(a.ctypes.data), a.nbytes, H2D), "hipMemcpy")

    @property
    def address(self) -> int:
        return int(self.ptr.value or 0)

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy_d2d(dst: DeviceArray, src: DeviceArray, nbytes: int):
    if nbytes > min(dst.nbytes, src.nbytes):
        raise ValueError("copy larger than the buffers")
    _ok(hip().hipMemcpy(dst.ptr, src.ptr, nbytes, D2D), "hipMemcpy")


def mem_info():
    """(free, total) bytes of device memory."""
    f, t = C.c_size_t(), C.c_size_t()
    _ok(hip().hipMemGetInfo(C.byref(f), C.byref(t)), "hipMemGetInfo")
    return f.value, t.value


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
