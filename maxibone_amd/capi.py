"""ctypes bindings to the C ABI of libmxb_diffusion (include/maxibone_diffusion.h,
include/maxibone_fields.h, include/maxibone_edt.h, include/maxibone_probe.h).

The library is loaded from the package directory, where ``maxibone_amd.build``
places it.  There is no fallback: if the library is missing, ``lib()`` raises.
"""
import ctypes
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmxb_diffusion.so")
HEADERS = [os.path.join(os.path.dirname(PKG), "include", h)
           for h in ("maxibone_diffusion.h", "maxibone_fields.h", "maxibone_edt.h", "maxibone_probe.h")]

OK, EINVAL, EDEVICE, ENOMEM, EIO = 0, 1, 2, 3, 4
PRECISION = {"fp32": 0, "bf16": 1, "fp16": 2, "fp32_fma": 3, "fp32_exact": 4, "fp32_split": 5}
K_ZPASS, K_YPASS, K_XPASS, K_OTHER, K_XYPASS = 0, 1, 2, 3, 4
NOALLOC_SCRATCH_BYTES = 20 << 20  # MXB_NOALLOC_SCRATCH_BYTES

_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_LIB = None


class DeviceVolume(ctypes.Structure):
    _fields_ = [("dev_mask", _vp), ("nz", _i64), ("ny", _i64), ("nx", _i64), ("kernel", _vp),
                ("kernel_size", _i64), ("dev_out_u16", _vp), ("dev_out_f32", _vp)]


class VolumeDesc(ctypes.Structure):
    _fields_ = [("mask", _vp), ("nz", _i64), ("ny", _i64), ("nx", _i64), ("kernel", _vp),
                ("kernel_size", _i64), ("output_u16", _vp), ("output_f32", _vp)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m maxibone_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "mxb_diffusion_u8_u16": [_vp, _i64, _i64, _i64, _vp, _i64, _i64, _vp, ctypes.c_int],
            "mxb_diffusion_u8_f32": [_vp, _i64, _i64, _i64, _vp, _i64, _i64, _vp, ctypes.c_int],
            "mxb_diffusion_ex": [_vp, _i64, _i64, _i64, _vp, _i64, _i64, ctypes.c_int, _vp, _vp,
                                 ctypes.c_int],
            "mxb_diffusion_chunked": [_vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64, _i64,
                                      ctypes.c_int, _vp, _vp, ctypes.c_int],
            "mxb_diffusion_on_disk": [ctypes.c_char_p, _vp, _i64, ctypes.c_char_p, _i64, _i64, _i64,
                                      _i64, _i64, _i64, _i64, ctypes.c_int],
            "mxb_diffusion_on_disk_ex": [ctypes.c_char_p, _vp, _i64, ctypes.c_char_p, _i64, _i64, _i64,
                                         _i64, _i64, _i64, _i64, ctypes.c_int, ctypes.c_int, ctypes.c_int],
            "mxb_diffusion_batched": [ctypes.POINTER(VolumeDesc), ctypes.c_int, _i64, ctypes.c_int,
                                      ctypes.c_int],
            "mxb_device_diffusion": [_vp, _i64, _i64, _i64, _vp, _i64, _i64, ctypes.c_int, _vp, _vp,
                                     _vp, _vp, _vp],
            "mxb_device_diffusion_noalloc": [_vp, _i64, _i64, _i64, _vp, _i64, _i64, ctypes.c_int, _vp,
                                             _vp, _vp],
            "mxb_set_kernel_timing": [ctypes.c_int],
            "mxb_kernel_times": [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(_i64)],
            "mxb_device_count": [],
            "mxb_batch_create": [ctypes.POINTER(DeviceVolume), ctypes.c_int, _i64, ctypes.c_int, _vp,
                                 ctypes.POINTER(_vp)],
            "mxb_batch_run": [_vp, ctypes.c_int],
            "mxb_batch_sync": [_vp],
            "mxb_batch_info": [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                               ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)],
            "mxb_batch_destroy": [_vp],
            "mxb_combine_fields": [_vp, _vp, _i64, _i64, _i64, _vp, _vp, ctypes.c_int],
            "mxb_device_combine_fields": [_vp, _vp, _i64, _i64, _i64, _vp, _vp, _vp, _vp],
            "mxb_field_histogram": [_vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64, _vp, _i64, _i64,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int],
            "mxb_device_field_histogram": [_vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64, _vp, _i64,
                                           _i64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, _vp, _vp],
            "mxb_axis_histograms": [_vp, _i64, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _i64, _i64,
                                    ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int],
            "mxb_device_axis_histograms": [_vp, _i64, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _i64,
                                           _i64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double,
                                           ctypes.c_double, _vp, _vp],
            "mxb_edt": [_vp, _i64, _i64, _i64, _vp, _vp, ctypes.c_int],
            "mxb_edt_field": [_vp, _i64, _i64, _i64, _vp, _vp, ctypes.c_int],
            "mxb_device_edt": [_vp, _i64, _i64, _i64, _vp, _vp, _vp, _vp],
            "mxb_device_edt_field": [_vp, _i64, _i64, _i64, _vp, _vp, _vp, _vp],
            "mxb_probe_hbm": [ctypes.c_int, _vp, _vp, _i64, ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_float)],
            "mxb_probe_mfma": [ctypes.c_int, _i64, _vp, ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_float),
                               ctypes.POINTER(ctypes.c_double)],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.mxb_state_bytes.argtypes = [ctypes.c_int, _i64, _i64, _i64]
        L.mxb_state_bytes.restype = _i64
        L.mxb_combine_workspace_bytes.argtypes = []
        L.mxb_combine_workspace_bytes.restype = _i64
        L.mxb_histogram_workspace_bytes.argtypes = [_i64, _i64, _i64]
        L.mxb_histogram_workspace_bytes.restype = _i64
        L.mxb_edt_workspace_bytes.argtypes = [_i64, _i64, _i64]
        L.mxb_edt_workspace_bytes.restype = _i64
        L.mxb_last_error.restype = ctypes.c_char_p
        L.mxb_version.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


class MxbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(rc):
    if rc != OK:
        msg = lib().mxb_last_error().decode()
        if rc == EINVAL:
            raise ValueError(msg)
        if rc == ENOMEM:
            raise MemoryError(msg)
        raise MxbError(rc, msg)


def declared_functions():
    """Function names declared in include/*.h."""
    names = set()
    for h in HEADERS:
        text = open(h).read()
        names |= set(re.findall(r"^\s*(?:const\s+)?[a-z0-9_]+\s+\*?\s*(mxb_[a-z0-9_]+)\s*\(", text,
                                flags=re.M))
    return sorted(names)


def device_count():
    return lib().mxb_device_count()


def _ptr(a):
    return None if a is None else a.ctypes.data


def diffusion(mask, kernel, reps, precision="fp32", u16=True, f32=False, verbose=0):
    """Host-buffer diffusion through mxb_diffusion_ex; returns (u16 or None, f32 or None)."""
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    k = np.ascontiguousarray(kernel, dtype=np.float32)
    q = np.empty(m.shape, np.uint16) if u16 else None
    f = np.empty(m.shape, np.float32) if f32 else None
    check(lib().mxb_diffusion_ex(m.ctypes.data, *m.shape, k.ctypes.data, k.shape[0], int(reps),
                                 PRECISION[precision], _ptr(q), _ptr(f), verbose))
    return q, f


def diffusion_chunked(mask, kernel, reps, chunk_z, temporal=1, precision="fp32", u16=True,
                      f32=False, verbose=0):
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    k = np.ascontiguousarray(kernel, dtype=np.float32)
    q = np.empty(m.shape, np.uint16) if u16 else None
    f = np.empty(m.shape, np.float32) if f32 else None
    check(lib().mxb_diffusion_chunked(m.ctypes.data, *m.shape, k.ctypes.data, k.shape[0], int(reps),
                                      int(chunk_z), int(temporal), PRECISION[precision], _ptr(q),
                                      _ptr(f), verbose))
    return q, f


ON_DISK_REFERENCE_MASK = 1


def diffusion_on_disk(input_file, kernel, output_file, total_shape, chunk_z, reps, precision="fp32",
                      reference_mask=False, verbose=0):
    """File-backed diffusion (mxb_diffusion_on_disk_ex): raw uint8 (z, y, x) in,
    raw uint16 out, z-chunks of ``chunk_z`` layers.  ``reference_mask``
    illuminates with (u == 1.0f) of each chunk's incoming state, the
    reference's on-disk rule."""
    k = np.ascontiguousarray(kernel, dtype=np.float32)
    tz, ty, tx = (int(v) for v in total_shape)
    flags = ON_DISK_REFERENCE_MASK if reference_mask else 0
    check(lib().mxb_diffusion_on_disk_ex(os.fsencode(input_file), k.ctypes.data, k.shape[0],
                                         os.fsencode(output_file), tz, ty, tx, int(chunk_z), ty, tx,
                                         int(reps), PRECISION[precision], flags, verbose))


def diffusion_batched(masks, kernels, reps, precision="fp32", u16=True, f32=False, verbose=0):
    ms = [np.ascontiguousarray(m, dtype=np.uint8) for m in masks]
    ks = [np.ascontiguousarray(k, dtype=np.float32) for k in kernels]
    qs = [np.empty(m.shape, np.uint16) if u16 else None for m in ms]
    fs = [np.empty(m.shape, np.float32) if f32 else None for m in ms]
    descs = (VolumeDesc * len(ms))()
    for i, (m, k, q, f) in enumerate(zip(ms, ks, qs, fs)):
        descs[i] = VolumeDesc(m.ctypes.data, *m.shape, k.ctypes.data, k.shape[0], _ptr(q), _ptr(f))
    check(lib().mxb_diffusion_batched(descs, len(ms), int(reps), PRECISION[precision], verbose))
    return qs, fs


def set_kernel_timing(enable):
    check(lib().mxb_set_kernel_timing(1 if enable else 0))


def kernel_times(cls):
    ms = ctypes.c_double()
    n = _i64()
    check(lib().mxb_kernel_times(cls, ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


class BatchPlan:
    """A device-resident batch (mxb_batch_create): volumes given as device
    pointers of the library's HIP runtime (maxibone_amd.hip buffers), kernels
    as numpy arrays."""

    def __init__(self, masks_dev, shapes, kernels, reps, outs_u16=None, outs_f32=None,
                 precision="fp32", stream=None):
        n = len(shapes)
        self._kernels = [np.ascontiguousarray(k, dtype=np.float32) for k in kernels]
        self._vols = (DeviceVolume * n)()
        for i in range(n):
            nz, ny, nx = shapes[i]
            self._vols[i] = DeviceVolume(masks_dev[i], nz, ny, nx, self._kernels[i].ctypes.data,
                                         self._kernels[i].shape[0],
                                         outs_u16[i] if outs_u16 else None,
                                         outs_f32[i] if outs_f32 else None)
        self._plan = _vp()
        check(lib().mxb_batch_create(self._vols, n, int(reps), PRECISION[precision], stream,
                                     ctypes.byref(self._plan)))

    def run(self, mode=0):
        check(lib().mxb_batch_run(self._plan, int(mode)))

    def sync(self):
        check(lib().mxb_batch_sync(self._plan))

    def info(self):
        p, g, s = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = ctypes.c_double()
        check(lib().mxb_batch_info(self._plan, ctypes.byref(p), ctypes.byref(g), ctypes.byref(s),
                                   ctypes.byref(ms)))
        return {"persistent": bool(p.value), "grid": g.value, "seg": s.value, "tuned_ms": ms.value}

    def close(self):
        if self._plan:
            lib().mxb_batch_destroy(self._plan)
            self._plan = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
