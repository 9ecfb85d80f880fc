"""Device buffers for the GPU tests through the HIP runtime that
libsilt_amd.so itself links (libamdhip64.so.7), so the tests' allocations and
the library's kernels share one runtime instance.  torch bundles a second HIP
runtime; mixing the two in one process leaves torch without a device."""
import ctypes

import numpy as np

import silt_amd as S  # noqa: F401  (loads libsilt_amd.so, and with it libamdhip64.so.7)

_H2D, _D2H = 1, 2
_hip = ctypes.CDLL("libamdhip64.so.7")
_hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
_hip.hipFree.argtypes = [ctypes.c_void_p]
_hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
_hip.hipDeviceSynchronize.argtypes = []


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class DeviceArray:
    """A device copy of a 1-D numpy array (same dtype and length)."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.dtype, self.size, self.nbytes = host.dtype, host.size, host.nbytes
        p = ctypes.c_void_p()
        _check(_hip.hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value
        _check(_hip.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, _H2D), "hipMemcpy H2D")

    def numpy(self):
        out = np.empty(self.size, self.dtype)
        _check(_hip.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            _hip.hipFree(self.ptr)
            self.ptr = None


def synchronize():
    _check(_hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
