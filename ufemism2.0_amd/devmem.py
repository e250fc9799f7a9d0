"""Device buffers for callers of the device-resident entry points
(ufe_heat_columns_device), through the HIP runtime that libufe_hip.so itself
links: a process holds one HIP runtime, so device memory comes from that one
and not from another framework's bundled copy."""

from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from . import _lib as L

_HOST_TO_DEVICE, _DEVICE_TO_HOST = 1, 2  # hipMemcpyKind
_hip = None


def hip():
    """The HIP runtime library the engine was linked against (loaded with it)."""
    global _hip
    if _hip is None:
        L.lib()
        last = None
        for name in ("/opt/rocm/lib/libamdhip64.so.7", "libamdhip64.so.7", "libamdhip64.so"):
            try:
                h = C.CDLL(name)
                break
            except OSError as e:
                last = e
        else:
            raise RuntimeError(f"HIP runtime not found: {last}")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipSetDevice.argtypes = [C.c_int]
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipSetDevice, h.hipDeviceSynchronize):
            f.restype = C.c_int
        _hip = h
    return _hip


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """One device allocation holding a 1-D array of a numpy dtype."""

    def __init__(self, count: int, dtype, device: int = 0):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        self.device = device
        h = hip()
        _check(h.hipSetDevice(device), "hipSetDevice")
        p = C.c_void_p()
        _check(h.hipMalloc(C.byref(p), max(self.nbytes, 8)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, device: int = 0) -> "DeviceArray":
        a = np.ascontiguousarray(a).reshape(-1)
        d = cls(a.size, a.dtype, device)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype).reshape(-1)
        if a.size != self.count:
            raise ValueError(f"size {a.size} != {self.count}")
        _check(hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, _HOST_TO_DEVICE), "hipMemcpy")

    def to_host(self) -> np.ndarray:
        out = np.empty(self.count, self.dtype)
        _check(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _DEVICE_TO_HOST), "hipMemcpy")
        return out

    def free(self):
        if self.ptr:
            _check(hip().hipFree(self.ptr), "hipFree")
            self.ptr = None

    def __del__(self):
        # at interpreter exit the runtime may already be torn down; the
        # process end releases the memory
        if sys.is_finalizing():
            return
        try:
            self.free()
        except Exception:
            pass
