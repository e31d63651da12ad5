"""Device buffers for the GPU tests, from the HIP runtime the library itself
uses (the libamdhip64 that libmaos_tomo.so links), called through ctypes.

Test helpers only.  PyTorch bundles a HIP runtime of its own; a process that
initialises the library's runtime first cannot then initialise torch's, so
the tests that need raw device memory take it from here instead."""
import ctypes as C

import numpy as np

from maos_amd import _abi

_hip = None


def hip():
    global _hip
    if _hip is None:
        _abi.lib()  # loads the library and, with it, its HIP runtime
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line and "/torch/" not in line:
                    path = line.split()[-1]
                    break
        if path is None:
            raise RuntimeError("the library's HIP runtime is not loaded")
        L = C.CDLL(path)
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipSetDevice.argtypes = [C.c_int]
        _hip = L
    return _hip


H2D, D2H, D2D = 1, 2, 3


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with HIP error %d" % (what, rc))


class DevArray:
    """A device buffer of a numpy dtype and shape (device 0)."""

    def __init__(self, shape, dtype):
        self.shape, self.dtype = tuple(np.atleast_1d(shape)), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _chk(hip().hipSetDevice(0), "hipSetDevice")
        _chk(hip().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def of(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        _chk(hip().hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), self.nbytes, H2D), "hipMemcpy")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _chk(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        _chk(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, D2H), "hipMemcpy")
        return out

    def copy(self):
        d = DevArray(self.shape, self.dtype)
        _chk(hip().hipMemcpy(C.c_void_p(d.ptr), C.c_void_p(self.ptr), self.nbytes, D2D), "hipMemcpy")
        return d

    def free(self):
        if self.ptr:
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def sync():
    _chk(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
