"""Minimal ctypes access to the HIP runtime that libcapaz_amd.so runs on
(device buffers, copies and events on a context's stream), so host code can
time kernels and stage data without bringing a second HIP runtime into the
process."""
from __future__ import annotations

import ctypes

from . import capaz

_hip = None

H2D, D2H, D2D = 1, 2, 3


def lib() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        capaz.load()  # resolves libamdhip64.so.7 for libcapaz_amd.so first
        h = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        h.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        h.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipEventRecord.argtypes = [vp, vp]
        h.hipEventSynchronize.argtypes = [vp]
        h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        h.hipEventDestroy.argtypes = [vp]
        h.hipGetErrorString.restype = ctypes.c_char_p
        _hip = h
    return _hip


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {lib().hipGetErrorString(rc).decode()}")


def copy(dst: int, src: int, nbytes: int) -> None:
    """Device-to-device copy (hipMemcpy; returns when it is done)."""
    _ok(lib().hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, D2D), "hipMemcpy")


class DeviceBuffer:
    """nbytes of device memory (hipMalloc), freed on close()."""

    def __init__(self, nbytes: int):
        self.ptr = ctypes.c_void_p()
        self.nbytes = nbytes
        _ok(lib().hipMalloc(ctypes.byref(self.ptr), nbytes), "hipMalloc")

    @property
    def address(self) -> int:
        return self.ptr.value or 0

    def upload(self, array) -> None:
        import numpy as np
        a = np.ascontiguousarray(array)
        assert a.nbytes <= self.nbytes
        _ok(lib().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D), "hipMemcpy")

    def download(self, count: int, dtype="float64"):
        import numpy as np
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _ok(lib().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, D2H), "hipMemcpy")
        return out

    def close(self) -> None:
        if self.ptr.value:
            lib().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.ev = ctypes.c_void_p()
        _ok(lib().hipEventCreate(ctypes.byref(self.ev)), "hipEventCreate")

    def record(self, stream: int) -> None:
        _ok(lib().hipEventRecord(self.ev, ctypes.c_void_p(stream)), "hipEventRecord")

    def synchronize(self) -> None:
        _ok(lib().hipEventSynchronize(self.ev), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        _ok(lib().hipEventElapsedTime(ctypes.byref(ms), self.ev, end.ev), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            lib().hipEventDestroy(self.ev)
        except Exception:
            pass
