"""GPU 2D histograms of the uint16 volumes: a drop-in for the reference's
``histograms`` module (src/pybind/histograms-pybind.cc:154-178, as imported by
src/processing_steps/1000_compute_histograms.py:18-20).

``axis_histograms(np_voxels, np_offset, np_x_bins, np_y_bins, np_z_bins,
np_r_bins, center, vrange, verbose=0)`` and ``field_histogram(np_voxels,
np_field, np_offset, np_bins, vrange, frange, verbose=0)`` add into the given
uint64 bin arrays in place, like the reference (which requires them to be
uint64, C-contiguous: ``noconvert``).  Bin semantics are the reference CPU
module's (src/lib/cpp/cpu/histograms.cc); bins it would index out of range
are skipped, and the voxel bin is floored (the reference's OpenACC axis
histogram rounds it, src/lib/cpp/gpu/histograms.cc:94).  Offsets in y and x
must be 0 (the reference's callers pass (z, 0, 0)).
"""
import numpy as np

from . import capi


def _bins(a, name):
    if not isinstance(a, np.ndarray) or a.dtype != np.uint64 or not a.flags.c_contiguous \
            or not a.flags.writeable or a.ndim != 2:
        raise TypeError(f"{name} must be a writable C-contiguous 2D uint64 array")
    return a


def _voxels(a, name, ndim=3):
    a = np.ascontiguousarray(a)
    if a.dtype != np.uint16:
        a = a.astype(np.uint16)  # forcecast, as the reference's np_array arguments
    if a.ndim != ndim:
        raise ValueError(f"{name} must be {ndim}-dimensional")
    return a


def _offset(np_offset):
    z, y, x = (int(v) for v in np_offset)
    if y != 0 or x != 0:
        raise ValueError("only z offsets are supported (y and x offsets must be 0)")
    return z


def axis_histograms(np_voxels, np_offset, np_x_bins, np_y_bins, np_z_bins, np_r_bins, center,
                    vrange, verbose=0):
    v = _voxels(np_voxels, "np_voxels")
    xb, yb, zb, rb = (_bins(b, n) for b, n in ((np_x_bins, "np_x_bins"), (np_y_bins, "np_y_bins"),
                                               (np_z_bins, "np_z_bins"), (np_r_bins, "np_r_bins")))
    nz, ny, nx = v.shape
    vb = xb.shape[1]
    if xb.shape[0] != nx or yb.shape[0] != ny or yb.shape[1] != vb or zb.shape[1] != vb or rb.shape[1] != vb:
        raise ValueError("bin arrays do not match the voxel chunk (x: (nx, vb), y: (ny, vb), z: (Nz, vb), "
                         "r: (Nr, vb))")
    cy, cx = (int(c) for c in center)
    vmin, vmax = (float(r) for r in vrange)
    capi.check(capi.lib().mxb_axis_histograms(v.ctypes.data, nz, ny, nx, _offset(np_offset), xb.ctypes.data,
                                              yb.ctypes.data, zb.ctypes.data, rb.ctypes.data, zb.shape[0],
                                              rb.shape[0], vb, cy, cx, vmin, vmax, int(verbose)))


def field_histogram(np_voxels, np_field, np_offset, np_bins, vrange, frange, verbose=0):
    v = _voxels(np_voxels, "np_voxels")
    f = _voxels(np_field, "np_field")
    b = _bins(np_bins, "np_bins")
    _offset(np_offset)
    vmin, vmax = (float(r) for r in vrange)
    fmin, fmax = (float(r) for r in frange)
    capi.check(capi.lib().mxb_field_histogram(v.ctypes.data, *v.shape, f.ctypes.data, *f.shape, b.ctypes.data,
                                              b.shape[0], b.shape[1], vmin, vmax, fmin, fmax, int(verbose)))
