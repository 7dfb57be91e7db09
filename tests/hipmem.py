"""Minimal device-memory plumbing for kernel-level tests (HIP runtime via
ctypes): allocate, copy, free.  No torch needed."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with HIP error %d" % (what, rc))


class DeviceArray:
    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.nbytes = max(8, self.host.nbytes)
        self.ptr = C.c_void_p()
        _check(hip().hipMalloc(C.byref(self.ptr), self.nbytes), "hipMalloc")
        self.upload()

    def upload(self):
        _check(hip().hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p), self.host.nbytes, 1), "hipMemcpy")

    def download(self):
        _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        _check(hip().hipMemcpy(self.host.ctypes.data_as(C.c_void_p), self.ptr, self.host.nbytes, 2), "hipMemcpy")
        return self.host

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


FRONT_DTYPE = np.dtype([("off", "<i8"), ("xoff", "<i8"), ("soff", "<i8"), ("n", "<i4"), ("fs", "<i4"),
                        ("poff", "<i4"), ("parent", "<i4"), ("emap", "<i4"), ("id", "<i4"), ("prec", "<i4"),
                        ("rsv", "<i4")])
assert FRONT_DTYPE.itemsize == 56
