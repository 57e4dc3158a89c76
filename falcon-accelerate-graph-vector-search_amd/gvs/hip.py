"""Minimal ctypes access to the HIP runtime that libgvs.so links
(libamdhip64.so.7): device buffers, streams and events for tests and the
benchmark, so no second runtime is loaded into the process."""
import ctypes as C

import numpy as np

from . import lib as _gvs  # noqa: F401  (ensures libgvs and its HIP runtime are loaded first)

hip = C.CDLL("libamdhip64.so.7")
_P = C.c_void_p
hip.hipMalloc.argtypes = [C.POINTER(_P), C.c_size_t]
hip.hipFree.argtypes = [_P]
hip.hipMemcpy.argtypes = [_P, _P, C.c_size_t, C.c_int]
hip.hipMemcpyAsync.argtypes = [_P, _P, C.c_size_t, C.c_int, _P]
hip.hipMemset.argtypes = [_P, C.c_int, C.c_size_t]
hip.hipStreamCreate.argtypes = [C.POINTER(_P)]
hip.hipStreamDestroy.argtypes = [_P]
hip.hipStreamSynchronize.argtypes = [_P]
hip.hipDeviceSynchronize.argtypes = []
hip.hipEventCreate.argtypes = [C.POINTER(_P)]
hip.hipEventDestroy.argtypes = [_P]
hip.hipEventRecord.argtypes = [_P, _P]
hip.hipEventSynchronize.argtypes = [_P]
hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), _P, _P]
hip.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
hip.hipGetErrorString.restype = C.c_char_p
hip.hipGetErrorString.argtypes = [C.c_int]
H2D, D2H, D2D = 1, 2, 3


def check(rc, what="hip"):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip.hipGetErrorString(rc).decode()}")


def device_count():
    n = C.c_int(0)
    rc = hip.hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


class DeviceArray:
    """A device buffer with a numpy dtype/shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (shape,)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = _P()
        check(hip.hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        check(hip.hipMemcpy(d.ptr, a.ctypes.data_as(_P), a.nbytes, H2D), "hipMemcpy H2D")
        return d

    def upload(self, a):
        """Copy a host array of the same byte size into this buffer."""
        a = np.ascontiguousarray(a)
        if a.nbytes != self.nbytes:
            raise ValueError(f"upload of {a.nbytes} bytes into a {self.nbytes}-byte buffer")
        check(hip.hipMemcpy(self.ptr, a.ctypes.data_as(_P), a.nbytes, H2D), "hipMemcpy H2D")

    @property
    def address(self):
        return self.ptr.value

    def to_host(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(hip.hipMemcpy(out.ctypes.data_as(_P), self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            hip.hipFree(self.ptr)
            self.ptr = _P()


class Stream:
    def __init__(self):
        self.s = _P()
        check(hip.hipStreamCreate(C.byref(self.s)), "hipStreamCreate")

    @property
    def handle(self):
        return self.s.value

    def synchronize(self):
        check(hip.hipStreamSynchronize(self.s), "hipStreamSynchronize")

    def __del__(self):
        if getattr(self, "s", None) is not None and self.s.value:
            hip.hipStreamDestroy(self.s)


class Event:
    def __init__(self):
        self.e = _P()
        check(hip.hipEventCreate(C.byref(self.e)), "hipEventCreate")

    def record(self, stream=None):
        check(hip.hipEventRecord(self.e, stream.s if stream is not None else None), "hipEventRecord")

    def synchronize(self):
        check(hip.hipEventSynchronize(self.e), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = C.c_float(0.0)
        check(hip.hipEventElapsedTime(C.byref(ms), self.e, end.e), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        if getattr(self, "e", None) is not None and self.e.value:
            hip.hipEventDestroy(self.e)


def synchronize():
    check(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
