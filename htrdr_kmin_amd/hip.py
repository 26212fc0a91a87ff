"""Minimal HIP runtime plumbing over ctypes: device buffers, streams, events.

It binds the very libamdhip64.so.7 that libhtgx.so is linked against (loaded
first, so the dynamic linker returns the same instance): one HIP runtime per
process. Used by tests and bench.py to keep accumulators resident in HBM and
to time kernels with HIP events on the stream they are launched on.
"""
import ctypes as C

from . import _ffi as F

_hip = None


def hip():
    global _hip
    if _hip is None:
        F.load()  # brings libamdhip64.so.7 in through libhtgx.so's DT_NEEDED
        h = C.CDLL("libamdhip64.so.7")
        sigs = {
            "hipMalloc": [C.POINTER(C.c_void_p), C.c_size_t],
            "hipFree": [C.c_void_p],
            "hipMemset": [C.c_void_p, C.c_int, C.c_size_t],
            "hipMemcpy": [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int],
            "hipDeviceSynchronize": [],
            "hipStreamCreate": [C.POINTER(C.c_void_p)],
            "hipStreamDestroy": [C.c_void_p],
            "hipStreamSynchronize": [C.c_void_p],
            "hipEventCreate": [C.POINTER(C.c_void_p)],
            "hipEventDestroy": [C.c_void_p],
            "hipEventRecord": [C.c_void_p, C.c_void_p],
            "hipEventSynchronize": [C.c_void_p],
            "hipEventElapsedTime": [C.POINTER(C.c_float), C.c_void_p, C.c_void_p],
            "hipGetDeviceCount": [C.POINTER(C.c_int)],
            "hipSetDevice": [C.c_int],
        }
        for name, args in sigs.items():
            fn = getattr(h, name)
            fn.restype = C.c_int
            fn.argtypes = args
        h.hipGetErrorString.restype = C.c_char_p
        h.hipGetErrorString.argtypes = [C.c_int]
        _hip = h
    return _hip


def ok(name, err):
    if err != 0:
        raise RuntimeError(f"{name}: {hip().hipGetErrorString(err).decode()}")


H2D, D2H = 1, 2


def device_count():
    n = C.c_int(0)
    err = hip().hipGetDeviceCount(C.byref(n))
    return n.value if err == 0 else 0


class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.ptr = C.c_void_p()
        ok("hipMalloc", hip().hipMalloc(C.byref(self.ptr), max(nbytes, 1)))

    def zero(self):
        ok("hipMemset", hip().hipMemset(self.ptr, 0, self.nbytes))

    def download(self, host_array):
        assert host_array.nbytes <= self.nbytes
        ok("hipMemcpy", hip().hipMemcpy(host_array.ctypes.data_as(C.c_void_p), self.ptr,
                                        host_array.nbytes, D2H))
        return host_array

    def upload(self, host_array):
        assert host_array.nbytes <= self.nbytes
        ok("hipMemcpy", hip().hipMemcpy(self.ptr, host_array.ctypes.data_as(C.c_void_p),
                                        host_array.nbytes, H2D))

    @property
    def address(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.h = C.c_void_p()
        ok("hipStreamCreate", hip().hipStreamCreate(C.byref(self.h)))

    def synchronize(self):
        ok("hipStreamSynchronize", hip().hipStreamSynchronize(self.h))

    def destroy(self):
        if self.h:
            hip().hipStreamDestroy(self.h)
            self.h = C.c_void_p()


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        ok("hipEventCreate", hip().hipEventCreate(C.byref(self.h)))

    def record(self, stream):
        ok("hipEventRecord", hip().hipEventRecord(self.h, stream.h if stream else None))

    def synchronize(self):
        ok("hipEventSynchronize", hip().hipEventSynchronize(self.h))

    def elapsed_ms(self, end):
        ms = C.c_float()
        ok("hipEventElapsedTime", hip().hipEventElapsedTime(C.byref(ms), self.h, end.h))
        return ms.value

    def destroy(self):
        if self.h:
            hip().hipEventDestroy(self.h)
            self.h = C.c_void_p()


def synchronize():
    ok("hipDeviceSynchronize", hip().hipDeviceSynchronize())
