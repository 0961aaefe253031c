"""Minimal ctypes access to the HIP runtime the library itself links (libamdhip64): events for
timing on a library stream, and raw device buffers for the pack/unpack entry points.

Loading the same runtime instance as libastaroth keeps device pointers and streams valid across
both; nothing here computes.
"""
import ctypes

import numpy as np

_H2D, _D2H = 1, 2


class Hip:
    def __init__(self):
        self.dll = ctypes.CDLL("libamdhip64.so")
        d = self.dll
        vp = ctypes.c_void_p
        d.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        d.hipEventRecord.argtypes = [vp, vp]
        d.hipEventSynchronize.argtypes = [vp]
        d.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        d.hipEventDestroy.argtypes = [vp]
        d.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        d.hipFree.argtypes = [vp]
        d.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        d.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
        d.hipDeviceSynchronize.argtypes = []

    # hipDeviceAttributeMultiprocessorCount in hip_runtime_api.h's hipDeviceAttribute_t
    _ATTR_MULTIPROCESSOR_COUNT = 63

    def cu_count(self, device: int = 0) -> int:
        n = ctypes.c_int()
        if self.dll.hipDeviceGetAttribute(ctypes.byref(n), self._ATTR_MULTIPROCESSOR_COUNT, device) != 0:
            raise RuntimeError("hipDeviceGetAttribute failed")
        return n.value

    # ---- events ----
    def event(self):
        ev = ctypes.c_void_p()
        if self.dll.hipEventCreate(ctypes.byref(ev)) != 0:
            raise RuntimeError("hipEventCreate failed")
        return ev

    def record(self, ev, stream):
        if self.dll.hipEventRecord(ev, stream) != 0:
            raise RuntimeError("hipEventRecord failed")

    def elapsed(self, a, b):
        ms = ctypes.c_float()
        self.dll.hipEventSynchronize(b)
        if self.dll.hipEventElapsedTime(ctypes.byref(ms), a, b) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return ms.value

    def sync(self):
        if self.dll.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")

    # ---- device buffers ----
    def malloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        if self.dll.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc(%d) failed" % nbytes)
        return p.value

    def free(self, ptr: int):
        if self.dll.hipFree(ctypes.c_void_p(ptr)) != 0:
            raise RuntimeError("hipFree failed")

    def memset(self, ptr: int, byte: int, nbytes: int):
        if self.dll.hipMemset(ctypes.c_void_p(ptr), byte, ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMemset failed")

    def to_device(self, ptr: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr)
        if self.dll.hipMemcpy(ctypes.c_void_p(ptr), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, _H2D) != 0:
            raise RuntimeError("hipMemcpy H2D failed")

    def to_host(self, ptr: int, count: int, dtype) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if self.dll.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), out.nbytes, _D2H) != 0:
            raise RuntimeError("hipMemcpy D2H failed")
        return out
