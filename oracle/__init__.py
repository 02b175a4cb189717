"""The oracle: a CPU restatement of maxibone's masked 3D diffusion.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import this package, and only as
the checker (or as the timed CPU baseline).  The product path
(``maxibone_amd``) never imports it and has no CPU fallback.

Contents
--------
* ``gauss_kernel(sigma)``: restates src/lib/py/helpers.py:376-409 (FP64 taps,
  radius = Python ``round(4*sigma)``, i.e. round-half-to-even, hlp:392;
  phi = exp(-0.5/sigma^2 * x^2) normalised by its FP64 sum, hlp:404-407).
* ``diffusion_f32`` / ``diffusion_f64`` / ``diffusion_chunked_f32`` /
  ``core_f32`` / ``quantise_u16``: ctypes wrappers over ``liboracle.so``
  (``diffusion_oracle.c``), which restates src/lib/cpp/cpu/diffusion.cc:31-60,
  191-196, 243-248, 295-302, 314-320, 339-374 and 376-468.
* ``edt``: the EDT field of src/processing_steps/0950_generate_edt.py
  (``edt_oracle.c`` for the distance transform; see ``oracle/edt.py``).
* ``ref_module()``: the reference's own CPU module compiled from its sources
  into ``oracle/_ref`` by ``oracle/build.py`` (None when not built).
"""
import ctypes
import importlib.util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64


def lib():
    """Load ``liboracle.so`` (built by ``oracle/build.py``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            from . import build as _build
            _build.build_oracle()
        L = ctypes.CDLL(path)
        L.oracle_u8_to_f32.argtypes = [_u8p, _f32p, _i64]
        L.oracle_quantise_u16.argtypes = [_f32p, _u16p, _i64]
        L.oracle_quantise_u8.argtypes = [_f32p, _u8p, _i64]
        L.oracle_core_f32.argtypes = [_f32p, _f32p, _f32p, _i64, _i64, _i64, ctypes.c_int, _i64]
        L.oracle_core_f64.argtypes = [_f64p, _f64p, _f64p, _i64, _i64, _i64, ctypes.c_int, _i64]
        L.oracle_diffusion_f32.argtypes = [_u8p, _i64, _i64, _i64, _f32p, _i64, _i64,
                                           ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_diffusion_f32.restype = ctypes.c_int
        L.oracle_diffusion_f64.argtypes = [_u8p, _i64, _i64, _i64, _f64p, _i64, _i64, _f64p]
        L.oracle_diffusion_f64.restype = ctypes.c_int
        L.oracle_diffusion_chunked_f32.argtypes = [_u8p, _i64, _i64, _i64, _f32p, _i64, _i64,
                                                   _i64, ctypes.c_int, _f32p]
        L.oracle_diffusion_chunked_f32.restype = ctypes.c_int
        L.oracle_edt_sq.argtypes = [_u8p, _i64, _i64, _i64,
                                    np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")]
        L.oracle_edt_sq.restype = ctypes.c_int
        _LIB = L
    return _LIB


def gauss_kernel(sigma):
    """FP64 Gaussian taps, restating src/lib/py/helpers.py:376-409."""
    radius = round(4.0 * sigma)  # hlp:392, Python's round-half-to-even
    sigma2 = sigma * sigma
    x = np.arange(-radius, radius + 1)
    phi_x = np.exp(-0.5 / sigma2 * x ** 2)
    return phi_x / phi_x.sum()  # hlp:404-407


def _mask_u8(mask):
    m = np.ascontiguousarray(mask)
    if m.dtype != np.uint8:
        m = m.astype(np.uint8)  # pybind forcecast semantics (dt:56-57)
    return m


def diffusion_f32(mask, kernel, reps):
    """Return (u float32, q uint16) of the reference CPU path (cpu:339-374)."""
    m = _mask_u8(mask)
    k = np.ascontiguousarray(kernel, dtype=np.float32)
    u = np.empty(m.shape, np.float32)
    q = np.empty(m.shape, np.uint16)
    rc = lib().oracle_diffusion_f32(m, *m.shape, k, k.shape[0], int(reps),
                                    u.ctypes.data, q.ctypes.data)
    if rc != 0:
        raise ValueError("oracle_diffusion_f32 rejected its arguments")
    return u, q


def diffusion_f64(mask, kernel, reps):
    """The same algorithm in FP64; ``kernel`` is promoted to float64 as given."""
    m = _mask_u8(mask)
    k = np.ascontiguousarray(kernel, dtype=np.float64)
    u = np.empty(m.shape, np.float64)
    rc = lib().oracle_diffusion_f64(m, *m.shape, k, k.shape[0], int(reps), u)
    if rc != 0:
        raise ValueError("oracle_diffusion_f64 rejected its arguments")
    return u


def diffusion_chunked_f32(mask, kernel, reps, chunk_z, quirk_mask=False):
    """z-chunked restatement of cpu:376-468 (true-mask unless quirk_mask)."""
    m = _mask_u8(mask)
    k = np.ascontiguousarray(kernel, dtype=np.float32)
    u = np.empty(m.shape, np.float32)
    rc = lib().oracle_diffusion_chunked_f32(m, *m.shape, k, k.shape[0], int(reps),
                                            int(chunk_z), int(bool(quirk_mask)), u)
    if rc != 0:
        raise ValueError("oracle_diffusion_chunked_f32 rejected its arguments")
    return u


def core_f32(f, kernel, dim):
    """One 1D correlation pass along ``dim`` (cpu:31-60)."""
    f = np.ascontiguousarray(f, dtype=np.float32)
    k = np.ascontiguousarray(kernel, dtype=np.float32)
    out = np.empty_like(f)
    lib().oracle_core_f32(f, k, out, *f.shape, int(dim), k.shape[0] // 2)
    return out


def core_f64(f, kernel, dim):
    f = np.ascontiguousarray(f, dtype=np.float64)
    k = np.ascontiguousarray(kernel, dtype=np.float64)
    out = np.empty_like(f)
    lib().oracle_core_f64(f, k, out, *f.shape, int(dim), k.shape[0] // 2)
    return out


def quantise_u16(u):
    """floor(u * 65535.0f) as uint16 (cpu:191-196)."""
    u = np.ascontiguousarray(u, dtype=np.float32)
    q = np.empty(u.shape, np.uint16)
    lib().oracle_quantise_u16(u, q, u.size)
    return q


def quantise_u8(u):
    u = np.ascontiguousarray(u, dtype=np.float32)
    q = np.empty(u.shape, np.uint8)
    lib().oracle_quantise_u8(u, q, u.size)
    return q


def ref_module(name="diffusion"):
    """A reference CPU module (diffusion, histograms, general) built from
    source into oracle/_ref, or None."""
    from .build import ref_module_path
    path = ref_module_path(name)
    if not os.path.exists(path):
        return None
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod
