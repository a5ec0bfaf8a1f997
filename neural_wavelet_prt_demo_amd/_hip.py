"""Minimal HIP runtime plumbing over ctypes: device buffers, copies, events.

Only what tests and the benchmark need to hand device pointers to libnwprt
and to time its launches.  The runtime is the ``libamdhip64.so.7`` that
libnwprt itself links, so both see the same device allocations.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

_hip = None

H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        _hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                        ctypes.c_int, ctypes.c_void_p]
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                             ctypes.c_void_p]
        _hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipStreamWaitEvent.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint]
        _hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
    return _hip


def check(err: int, what: str) -> None:
    if err != 0:
        msg = hip().hipGetErrorString(err).decode()
        raise RuntimeError(f"{what}: HIP error {err} ({msg})")


def device_count() -> int:
    n = ctypes.c_int(0)
    err = hip().hipGetDeviceCount(ctypes.byref(n))
    return n.value if err == 0 else 0


def synchronize() -> None:
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def stream_synchronize(stream: Optional[int]) -> None:
    check(hip().hipStreamSynchronize(ctypes.c_void_p(stream or 0)), "hipStreamSynchronize")


class DeviceBuffer:
    """A hipMalloc'd byte buffer; freed when the object is collected."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes, "upload larger than buffer"
        check(hip().hipMemcpy(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p),
                              a.nbytes, H2D), "hipMemcpy H2D")

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes, "download larger than buffer"
        check(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                              out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def zero(self) -> None:
        check(hip().hipMemset(ctypes.c_void_p(self.ptr), 0, self.nbytes), "hipMemset")

    def copy_from(self, src: "DeviceBuffer", nbytes: int, stream: Optional[int] = None) -> None:
        """Device-to-device copy of ``nbytes`` on ``stream`` (a blit kernel)."""
        assert nbytes <= self.nbytes and nbytes <= src.nbytes, "copy larger than buffer"
        check(hip().hipMemcpyAsync(ctypes.c_void_p(self.ptr), ctypes.c_void_p(src.ptr),
                                   ctypes.c_size_t(nbytes), D2D, ctypes.c_void_p(stream or 0)),
              "hipMemcpyAsync D2D")

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(hip().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.ev = e.value

    def record(self, stream: Optional[int]) -> None:
        check(hip().hipEventRecord(ctypes.c_void_p(self.ev), ctypes.c_void_p(stream or 0)),
              "hipEventRecord")

    def synchronize(self) -> None:
        check(hip().hipEventSynchronize(ctypes.c_void_p(self.ev)), "hipEventSynchronize")

    def wait(self, stream: Optional[int]) -> None:
        """Make later work on ``stream`` wait for this event's last record."""
        check(hip().hipStreamWaitEvent(ctypes.c_void_p(stream or 0), ctypes.c_void_p(self.ev), 0),
              "hipStreamWaitEvent")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0)
        check(hip().hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(self.ev),
                                        ctypes.c_void_p(end.ev)), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            hip().hipEventDestroy(ctypes.c_void_p(self.ev))
        except Exception:
            pass


class Stream:
    """A HIP stream at the default priority; destroyed when collected."""

    def __init__(self):
        st = ctypes.c_void_p()
        check(hip().hipStreamCreate(ctypes.byref(st)), "hipStreamCreate")
        self.handle = st.value

    def synchronize(self) -> None:
        stream_synchronize(self.handle)

    def __del__(self):
        try:
            if self.handle:
                hip().hipStreamDestroy(ctypes.c_void_p(self.handle))
                self.handle = None
        except Exception:
            pass
