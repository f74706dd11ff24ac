"""Minimal ctypes binding of the HIP runtime that libgpunb.so links
(libamdhip64.so.7 from /opt/rocm): device buffers and events for callers of
the device-resident entry points (gpunb_regf_dev) and for timing kernels on
the library's own stream.  Using the same runtime instance as the library
keeps pointers and streams valid across the two.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import gpunb

_rt = None
H2D, D2H, D2D = 1, 2, 3


def rt():
    global _rt
    if _rt is None:
        gpunb.load()  # the library pulls in its runtime first
        name = "libamdhip64.so.7"
        try:
            lib = ctypes.CDLL(name)
        except OSError:
            lib = ctypes.CDLL(os.path.join("/opt/rocm/lib", name))
        vp = ctypes.c_void_p
        lib.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        lib.hipFree.argtypes = [vp]
        lib.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        lib.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
        lib.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        lib.hipEventRecord.argtypes = [vp, vp]
        lib.hipEventSynchronize.argtypes = [vp]
        lib.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp,
                                            vp]
        lib.hipEventDestroy.argtypes = [vp]
        lib.hipStreamSynchronize.argtypes = [vp]
        lib.hipDeviceSynchronize.argtypes = []
        lib.hipGetErrorString.restype = ctypes.c_char_p
        lib.hipGetErrorString.argtypes = [ctypes.c_int]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemset",
                  "hipEventCreate", "hipEventRecord", "hipEventSynchronize",
                  "hipEventElapsedTime", "hipEventDestroy",
                  "hipStreamSynchronize", "hipDeviceSynchronize"):
            getattr(lib, f).restype = ctypes.c_int
        _rt = lib
    return _rt


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {rt().hipGetErrorString(rc).decode()}")


class DeviceArray:
    """A device allocation with a numpy dtype/shape view for copies."""

    def __init__(self, shape, dtype):
        self.shape = tuple(np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(rt().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        check(rt().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy")
        return d

    def zero(self):
        check(rt().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")
        return self

    def copy_from(self, other: "DeviceArray"):
        """Device-to-device copy of another array of the same size, complete
        on return (a device-to-device hipMemcpy may return before the copy
        is done, and other streams are not ordered after it)."""
        assert other.nbytes == self.nbytes
        check(rt().hipMemcpy(self.ptr, other.ptr, self.nbytes, D2D),
              "hipMemcpy")
        check(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")
        return self

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        check(rt().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H),
              "hipMemcpy")
        return out

    def free(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(rt().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.ev = e.value

    def record(self, stream: int):
        check(rt().hipEventRecord(self.ev, stream), "hipEventRecord")

    def synchronize(self):
        check(rt().hipEventSynchronize(self.ev), "hipEventSynchronize")

    def elapsed_ms(self, later: "Event") -> float:
        t = ctypes.c_float()
        check(rt().hipEventElapsedTime(ctypes.byref(t), self.ev, later.ev),
              "hipEventElapsedTime")
        return float(t.value)


def device_synchronize():
    check(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")
