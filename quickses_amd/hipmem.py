"""Device-memory plumbing over the HIP runtime that libquickses.so links (ctypes).

Used by the benchmark and the GPU tests to place inputs in HBM before calling the
device-resident entry point; no kernels here.
"""
import ctypes

import numpy as np

from . import lib as _qs_lib

_hip = None
H2D, D2H, D2D = 1, 2, 3


def hip():
    """The HIP runtime already loaded by libquickses.so (same soname, same instance)."""
    global _hip
    if _hip is None:
        _qs_lib()
        h = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        h.hipGetErrorString.restype = ctypes.c_char_p
        h.hipHostMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t, ctypes.c_uint]
        h.hipHostFree.argtypes = [vp]
        _hip = h
    return _hip


def _ck(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed: %s" % (what, hip().hipGetErrorString(rc).decode()))


def device_count():
    n = ctypes.c_int(0)
    rc = hip().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    _ck(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A device allocation holding a copy of a numpy array."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.shape, self.dtype, self.nbytes = host.shape, host.dtype, host.nbytes
        p = ctypes.c_void_p()
        _ck(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value
        _ck(hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def alloc(nbytes):
    """Uninitialised device allocation; returns its address (free with free())."""
    p = ctypes.c_void_p()
    _ck(hip().hipMalloc(ctypes.byref(p), max(int(nbytes), 1)), "hipMalloc")
    return p.value


def free(ptr):
    if ptr:
        hip().hipFree(ctypes.c_void_p(ptr))


def download(ptr, shape, dtype):
    out = np.empty(shape, dtype=dtype)
    if out.nbytes:
        _ck(hip().hipMemcpy(out.ctypes.data, ptr, out.nbytes, D2H), "hipMemcpy D2H")
    return out


def upload(ptr, host):
    """Copy a numpy array into an existing device allocation (synchronous)."""
    host = np.ascontiguousarray(host)
    if host.nbytes:
        _ck(hip().hipMemcpy(ptr, host.ctypes.data, host.nbytes, H2D), "hipMemcpy H2D")


def d2h_rate(nbytes=1 << 30, reps=3):
    """GB/s of a device-to-host copy into pinned host memory (the link's rate for result copies)."""
    import time

    d = alloc(nbytes)
    h = ctypes.c_void_p()
    _ck(hip().hipHostMalloc(ctypes.byref(h), nbytes, 0), "hipHostMalloc")
    try:
        _ck(hip().hipMemcpy(h.value, d, nbytes, D2H), "hipMemcpy D2H")
        best = None
        for _ in range(reps):
            t = time.perf_counter()
            _ck(hip().hipMemcpy(h.value, d, nbytes, D2H), "hipMemcpy D2H")
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
    finally:
        hip().hipHostFree(h)
        free(d)
    return nbytes / best / 1e9
