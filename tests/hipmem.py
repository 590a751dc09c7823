"""Device buffers for GPU tests through the system HIP runtime (the one
libpra_amd.so links), so tests can hand device pointers to the C ABI."""
import ctypes
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise OSError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return _hip


def sync():
    """every piece of device work so far has completed (null-stream copies and
    fills are not ordered with the non-blocking streams of the library)"""
    _chk(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")


class DeviceBuffer:
    """nbytes of device memory, freed with the object"""

    def __init__(self, nbytes):
        self.nbytes = max(8, int(nbytes))
        p = ctypes.c_void_p()
        _chk(hip().hipMalloc(ctypes.byref(p), self.nbytes), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            _chk(hip().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        sync()
        return b

    def fill(self, byte):
        _chk(hip().hipMemset(self.ptr, int(byte), self.nbytes), "hipMemset")
        sync()

    def to_array(self, dtype=np.uint64, count=None):
        count = self.nbytes // np.dtype(dtype).itemsize if count is None else count
        out = np.empty(count, dtype=dtype)
        sync()
        if out.nbytes:
            _chk(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def close(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
