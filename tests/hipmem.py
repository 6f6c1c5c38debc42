"""Device memory for tests that call entry points taking device pointers:
hipMalloc / hipMemcpy / hipFree of the HIP runtime through ctypes (no torch)."""
import ctypes as C

import numpy as np

_hip = None
H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip():
    global _hip
    if _hip is None:
        L = C.CDLL("libamdhip64.so")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        _hip = L
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """A device copy of a host array (or an uninitialised buffer of nbytes)."""

    def __init__(self, host=None, nbytes=None):
        self.nbytes = int(host.nbytes if host is not None else nbytes)
        p = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(p), max(self.nbytes, 8)), "hipMalloc")
        self.p = p
        if host is not None:
            h = np.ascontiguousarray(host)
            _ok(hip().hipMemcpy(self.p, h.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    @property
    def ptr(self):
        return self.p.value

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        _ok(hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.p:
            _ok(hip().hipFree(self.p), "hipFree")
            self.p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
