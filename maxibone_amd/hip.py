"""Minimal HIP runtime access (device memory, streams, events) through ctypes.

It binds the same `libamdhip64.so.7` that libmxb_diffusion.so links, so
buffers and streams made here belong to the library's runtime instance.
(torch ships its own HIP runtime under a different soname; mixing the two in
one process lets only the first to initialise see the GPU, so the bench and
the GPU tests use this module instead of torch.)
"""
import ctypes

import numpy as np

from . import capi

_HIP = None
H2D, D2H, D2D = 1, 2, 3


def rt():
    global _HIP
    if _HIP is None:
        capi.lib()  # load the library first so that its runtime is the one bound here
        H = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        H.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        H.hipFree.argtypes = [vp]
        H.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        H.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        H.hipMemcpy2D.argtypes = [vp, sz, vp, sz, sz, sz, ctypes.c_int]
        H.hipDeviceSynchronize.argtypes = []
        H.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        H.hipStreamDestroy.argtypes = [vp]
        H.hipStreamSynchronize.argtypes = [vp]
        H.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        H.hipEventDestroy.argtypes = [vp]
        H.hipEventRecord.argtypes = [vp, vp]
        H.hipEventSynchronize.argtypes = [vp]
        H.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        H.hipGetErrorString.argtypes = [ctypes.c_int]
        H.hipGetErrorString.restype = ctypes.c_char_p
        _HIP = H
    return _HIP


def check(err):
    if err != 0:
        raise RuntimeError(f"HIP error {err}: {rt().hipGetErrorString(err).decode()}")


class DeviceBuffer:
    """A hipMalloc'd buffer; `ptr` is its device address."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        if self.nbytes < 0:
            raise ValueError(f"negative buffer size {self.nbytes}")
        p = ctypes.c_void_p()
        check(rt().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(rt().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, H2D))
        return b

    def to_array(self, dtype, shape):
        out = np.empty(shape, dtype)
        check(rt().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H))
        return out

    def box(self, dtype, shape, origin, size):
        """Copy the box [origin, origin + size) of the (nz, ny, nx) C-order
        volume of `dtype` held in this buffer to a new host array."""
        dtype = np.dtype(dtype)
        nz, ny, nx = (int(v) for v in shape)
        z0, y0, x0 = (int(v) for v in origin)
        bz, by, bx = (int(v) for v in size)
        if not (0 <= z0 and z0 + bz <= nz and 0 <= y0 and y0 + by <= ny and 0 <= x0 and x0 + bx <= nx):
            raise ValueError(f"box {origin}+{size} outside the volume {shape}")
        if nz * ny * nx * dtype.itemsize > self.nbytes:
            raise ValueError("volume larger than the buffer")
        out = np.empty((bz, by, bx), dtype)
        e = dtype.itemsize
        for z in range(bz):
            src = self.ptr + (((z0 + z) * ny + y0) * nx + x0) * e
            check(rt().hipMemcpy2D(out[z].ctypes.data, bx * e, src, nx * e, bx * e, by, D2H))
        return out

    def fill_bytes(self, value):
        check(rt().hipMemset(self.ptr, int(value), self.nbytes))

    def free(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        check(rt().hipStreamCreate(ctypes.byref(s)))
        self.handle = s.value

    def synchronize(self):
        check(rt().hipStreamSynchronize(self.handle))

    def __del__(self):
        try:
            rt().hipStreamDestroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(rt().hipEventCreate(ctypes.byref(e)))
        self.handle = e.value

    def record(self, stream):
        check(rt().hipEventRecord(self.handle, stream.handle if stream else None))

    def elapsed_ms(self, later):
        check(rt().hipEventSynchronize(later.handle))
        ms = ctypes.c_float()
        check(rt().hipEventElapsedTime(ctypes.byref(ms), self.handle, later.handle))
        return ms.value


def synchronize():
    check(rt().hipDeviceSynchronize())
