"""Device arrays, streams and events over the HIP runtime librtgi itself links (libamdhip64 of the ROCm install), for
the entry points that take device memory (rtgi_world_gather_async). A second HIP runtime in the same process (the one
a PyTorch wheel bundles) does not see the GPUs once librtgi's has opened them, so device memory for librtgi comes from
here."""
import ctypes as C

import numpy as np

from ._lib import lib

_H2D, _D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost
_hip = None


def runtime():
    """The HIP runtime library librtgi loaded (resolved by its soname, so the same instance)."""
    global _hip
    if _hip is None:
        lib()  # librtgi first: its runtime is the one the process uses
        h = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        for name, args in (("hipMalloc", [C.POINTER(C.c_void_p), C.c_size_t]), ("hipFree", [C.c_void_p]),
                           ("hipMemcpy", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
                           ("hipStreamCreate", [C.POINTER(C.c_void_p)]), ("hipStreamDestroy", [C.c_void_p]),
                           ("hipStreamSynchronize", [C.c_void_p]), ("hipEventCreate", [C.POINTER(C.c_void_p)]),
                           ("hipEventDestroy", [C.c_void_p]), ("hipEventRecord", [C.c_void_p, C.c_void_p]),
                           ("hipEventSynchronize", [C.c_void_p]),
                           ("hipEventElapsedTime", [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]),
                           ("hipGetErrorString", [C.c_int])):
            f = getattr(h, name)
            f.argtypes = args
            f.restype = C.c_char_p if name == "hipGetErrorString" else C.c_int
        _hip = h
    return _hip


def _check(e, what):
    if e != 0:
        raise RuntimeError(f"{what}: {runtime().hipGetErrorString(e).decode()}")


class DeviceArray:
    """A float32 / uint32 array in device memory: upload from and download to numpy."""

    def __init__(self, shape, dtype=np.float32):
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(runtime().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _check(runtime().hipMemcpy(d.ptr, a.ctypes.data, d.nbytes, _H2D), "hipMemcpy")
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _check(runtime().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H), "hipMemcpy")
        return out

    def free(self):
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:  # interpreter shutdown
            pass


class Stream:
    """A HIP stream (default priority); .handle is what the entry points take."""

    def __init__(self):
        s = C.c_void_p()
        _check(runtime().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        _check(runtime().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            runtime().hipStreamDestroy(self.handle)
            self.handle = None


class Event:
    """A timing event: record(stream), then elapsed_ms(other) after both completed."""

    def __init__(self):
        e = C.c_void_p()
        _check(runtime().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream):
        _check(runtime().hipEventRecord(self.handle, stream.handle if isinstance(stream, Stream) else stream),
               "hipEventRecord")

    def elapsed_ms(self, end):
        _check(runtime().hipEventSynchronize(end.handle), "hipEventSynchronize")
        ms = C.c_float()
        _check(runtime().hipEventElapsedTime(C.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return ms.value

    def destroy(self):
        if self.handle:
            runtime().hipEventDestroy(self.handle)
            self.handle = None
