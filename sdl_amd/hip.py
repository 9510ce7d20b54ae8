"""Minimal ctypes access to the HIP runtime that libsdl_amd.so links
(libamdhip64.so.7): device buffers, copies and events for tests and the
benchmark.  Plumbing only."""
import ctypes

import numpy as np

from . import runtime

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        runtime.lib()  # load the product library first so both share one runtime
        L = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        L.hipEventDestroy.argtypes = [vp]
        L.hipEventRecord.argtypes = [vp, vp]
        L.hipEventSynchronize.argtypes = [vp]
        L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        L.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _hip = L
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


def device_count():
    n = ctypes.c_int(0)
    if hip().hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


class DeviceArray:
    """A device copy of a contiguous numpy array."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.nbytes = host.nbytes
        self.dtype, self.shape = host.dtype, host.shape
        self.ptr = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 4)), "hipMalloc")
        if self.nbytes:
            _ok(hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D), "hipMemcpy")

    @classmethod
    def empty(cls, n, dtype):
        """An uninitialised device array of n elements."""
        a = cls.__new__(cls)
        a.dtype, a.shape = np.dtype(dtype), (int(n),)
        a.nbytes = int(n) * a.dtype.itemsize
        a.ptr = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(a.ptr), max(a.nbytes, 4)), "hipMalloc")
        return a

    @property
    def address(self):
        return self.ptr.value

    def upload(self, host):
        """Copies a host array of at most nbytes bytes to the start of the array."""
        host = np.ascontiguousarray(host)
        if host.nbytes > self.nbytes:
            raise ValueError("host array larger than the device array")
        if host.nbytes:
            _ok(hip().hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, H2D), "hipMemcpy")

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.h = ctypes.c_void_p()
        _ok(hip().hipEventCreate(ctypes.byref(self.h)), "hipEventCreate")

    def record(self, stream):
        _ok(hip().hipEventRecord(self.h, ctypes.c_void_p(stream)), "hipEventRecord")

    def elapsed_ms(self, end):
        _ok(hip().hipEventSynchronize(end.h), "hipEventSynchronize")
        ms = ctypes.c_float(0)
        _ok(hip().hipEventElapsedTime(ctypes.byref(ms), self.h, end.h), "hipEventElapsedTime")
        return ms.value


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
