"""Device buffers for GPU tests through the HIP runtime the library itself
links (/opt/rocm/lib/libamdhip64.so.7): one runtime per process.  (torch's
wheel carries its own copy of the HIP runtime, which does not see the device
once this one owns it, so the tests do not use torch for device memory.)"""
import ctypes
import os

import numpy as np

_PATHS = ["/opt/rocm/lib/libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so", "libamdhip64.so.7"]
_hip = None
for _p in _PATHS:
    if _p.startswith("/") and not os.path.exists(_p):
        continue
    try:
        _hip = ctypes.CDLL(_p)
        break
    except OSError:
        continue

H2D, D2H = 1, 2


class DeviceArray:
    """n doubles of device memory on device 0."""

    def __init__(self, n):
        if _hip is None:
            raise RuntimeError("HIP runtime not found")
        self.n = int(n)
        self.ptr = ctypes.c_void_p()
        if _hip.hipSetDevice(0) != 0 or _hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(8 * self.n)) != 0:
            raise RuntimeError("hipMalloc failed")

    def upload(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == self.n
        rc = _hip.hipMemcpy(self.ptr, x.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(8 * self.n), H2D)
        assert rc == 0, rc

    def download(self):
        y = np.empty(self.n)
        rc = _hip.hipMemcpy(y.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(8 * self.n), D2H)
        assert rc == 0, rc
        return y

    @property
    def address(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        self.free()
