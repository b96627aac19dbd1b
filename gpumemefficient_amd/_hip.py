"""Device buffers for callers that hand the library HBM-resident data
(wf_stencil_device): hipMalloc / hipMemcpy / hipFree of the HIP runtime that
libwfdp.so itself links, through ctypes."""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os

import numpy as np

_hip = None
H2D, D2H = 1, 2


def _load() -> C.CDLL:
    global _hip
    if _hip is None:
        from . import _lib
        _lib.load()  # the runtime libwfdp.so links, already mapped
        name = None
        for cand in ("libamdhip64.so", "libamdhip64.so.7", "libamdhip64.so.6"):
            try:
                _hip = C.CDLL(cand)
                name = cand
                break
            except OSError:
                continue
        if _hip is None:
            path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
            _hip = C.CDLL(path)
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipSetDevice.argtypes = [C.c_int]
        del name
    return _hip


class DeviceBuffer:
    """An int32 array in HBM: DeviceBuffer.from_array(host) uploads it,
    .to_array() downloads it, .ptr is the device address."""

    def __init__(self, shape, device: int = 0):
        h = _load()
        self.shape = tuple(int(s) for s in shape)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * 4
        self.ptr = 0
        if h.hipSetDevice(device) != 0:
            raise RuntimeError("hipSetDevice failed")
        p = C.c_void_p()
        if h.hipMalloc(C.byref(p), max(self.nbytes, 4)) != 0:
            raise MemoryError(f"hipMalloc of {self.nbytes} bytes failed")
        self.ptr = int(p.value)

    @classmethod
    def from_array(cls, a: np.ndarray, device: int = 0) -> "DeviceBuffer":
        a = np.ascontiguousarray(a, dtype=np.int32)
        d = cls(a.shape, device)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.int32)
        assert a.nbytes == self.nbytes
        if _load().hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), self.nbytes, H2D) != 0:
            raise RuntimeError("hipMemcpy H2D failed")

    def to_array(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.int32)
        if _load().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, D2H) != 0:
            raise RuntimeError("hipMemcpy D2H failed")
        return out

    def free(self) -> None:
        if self.ptr:
            _load().hipFree(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
