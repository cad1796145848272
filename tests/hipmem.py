"""Device buffers for tests that call entry points taking device pointers.

Uses the HIP runtime the library itself links (libamdhip64.so.7, already
loaded with libgdiamond_hip.so, so both share one runtime instance): plain
hipMalloc / hipMemcpy / hipFree through ctypes. Test infrastructure only."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        import gdiamond_amd
        gdiamond_amd.lib()  # loads the runtime the library links
        H = C.CDLL("libamdhip64.so.7")
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipDeviceSynchronize.argtypes = []
        H.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        H.hipStreamDestroy.argtypes = [C.c_void_p]
        H.hipStreamSynchronize.argtypes = [C.c_void_p]
        for f in (H.hipMalloc, H.hipFree, H.hipMemcpy, H.hipDeviceSynchronize, H.hipStreamCreate,
                  H.hipStreamDestroy, H.hipStreamSynchronize):
            f.restype = C.c_int
        _hip = H
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """A device copy of a contiguous numpy array (same dtype and shape)."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.shape, self.dtype, self.nbytes = host.shape, host.dtype, host.nbytes
        p = C.c_void_p()
        rc = hip().hipMalloc(C.byref(p), max(1, self.nbytes))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed: {rc}")
        self.ptr = p.value
        self.upload(host)

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.nbytes == self.nbytes
        rc = hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed: {rc}")

    def download(self):
        out = np.empty(self.shape, self.dtype)
        rc = hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed: {rc}")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
