"""Field combination on the GPU: the counterpart of the numpy block of
src/processing_steps/0975_combine_fields.py:41-62.

``combine_fields(gauss, edt, cylinder_mask=None)`` returns the uint16
combined field exactly as the script computes it (float32 arithmetic in the
script's order, the C++ ``to_int`` of src/lib/cpp/include/general.hh:96-118,
cylinder mask applied before and after the conversion).  Both inputs are
uint16 (z, y, x) volumes of one shape (0975 loads the saved uint16 fields);
``cylinder_mask`` is the 2D (ny, nx) mask, by default
``generate_cylinder_mask(ny, nx)`` (0975:39 calls it with nx for square
slices).
"""
import numpy as np

from . import capi
from .helpers import generate_cylinder_mask


def combine_fields(gauss, edt, cylinder_mask=None, verbose=0):
    g = np.ascontiguousarray(gauss)
    e = np.ascontiguousarray(edt)
    if g.dtype != np.uint16 or e.dtype != np.uint16:
        raise TypeError("combine_fields expects uint16 fields")
    if g.ndim != 3 or g.shape != e.shape:
        raise ValueError(f"fields must be 3D and of one shape, got {g.shape} and {e.shape}")
    nz, ny, nx = g.shape
    cyl = generate_cylinder_mask(ny, nx) if cylinder_mask is None else np.asarray(cylinder_mask)
    if cyl.shape != (ny, nx):
        raise ValueError(f"cylinder mask must be ({ny}, {nx}), got {cyl.shape}")
    c = np.ascontiguousarray(cyl != 0, dtype=np.uint8)
    out = np.empty(g.shape, np.uint16)
    capi.check(capi.lib().mxb_combine_fields(g.ctypes.data, e.ctypes.data, nz, ny, nx, c.ctypes.data,
                                             out.ctypes.data, int(verbose)))
    return out
