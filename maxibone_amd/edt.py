"""The implant EDT field on the GPU: the counterpart of
src/processing_steps/0950_generate_edt.py.

* ``edt(data, parallel=None)``: the Euclidean distance transform of a binary
  volume, as 0950:49 calls ``edt.edt(implant_mask, parallel=n_cores)`` (the
  external ``edt`` package with its defaults: anisotropy (1, 1, 1),
  black_border=False).  Every nonzero voxel gets the float32 distance to the
  nearest zero voxel; zero voxels get 0; +inf everywhere when the volume has
  no zero voxel.  ``parallel`` is accepted for drop-in use and ignored.
* ``edt_sq(data)``: the exact squared distances (uint32; ``EDT_NONE`` where
  no zero voxel exists).
* ``generate_edt_field(implant_solid, cylinder_mask=None)``: 0950:41-66:
  negate the implant mask, EDT, multiply by the cylinder mask, ``to_int`` to
  uint16 (the C++ normalized_convert of general.hh:96-118), multiply by the
  cylinder mask again.  ``cylinder_mask`` defaults to
  ``generate_cylinder_mask(ny, nx)`` (0950:51 uses generate_cylinder_mask(nx)
  for square slices).

Command line (0950's steps on a .npy implant mask, the reference reads it
from HDF5): ``python -m maxibone_amd.edt implant.npy --output edt.npy``.
Each axis may hold at most ``MAX_AXIS`` voxels.
"""
import argparse

import numpy as np

from . import capi
from .helpers import generate_cylinder_mask

EDT_NONE = 0xFFFFFFFF
MAX_AXIS = 4096


def _binary(data):
    d = np.asarray(data)
    if d.ndim != 3:
        raise ValueError(f"edt expects a 3D volume, got shape {d.shape}")
    if d.dtype != np.uint8 and d.dtype != np.bool_:
        d = d != 0
    return np.ascontiguousarray(d.view(np.uint8) if d.dtype == np.bool_ else d)


def edt(data, parallel=None, verbose=0):
    del parallel  # the GPU form has no host thread count
    d = _binary(data)
    out = np.empty(d.shape, np.float32)
    if d.size:
        capi.check(capi.lib().mxb_edt(d.ctypes.data, *d.shape, None, out.ctypes.data, int(verbose)))
    return out


def edt_sq(data, verbose=0):
    d = _binary(data)
    out = np.empty(d.shape, np.uint32)
    if d.size:
        capi.check(capi.lib().mxb_edt(d.ctypes.data, *d.shape, out.ctypes.data, None, int(verbose)))
    return out


def generate_edt_field(implant_solid, cylinder_mask=None, verbose=0):
    m = _binary(implant_solid)
    nz, ny, nx = m.shape
    cyl = generate_cylinder_mask(ny, nx) if cylinder_mask is None else np.asarray(cylinder_mask)
    if cyl.shape != (ny, nx):
        raise ValueError(f"cylinder mask must be ({ny}, {nx}), got {cyl.shape}")
    c = np.ascontiguousarray(cyl != 0, dtype=np.uint8)
    out = np.empty(m.shape, np.uint16)
    if m.size:
        capi.check(capi.lib().mxb_edt_field(m.ctypes.data, nz, ny, nx, c.ctypes.data, out.ctypes.data,
                                            int(verbose)))
    return out


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("implant", help=".npy implant_solid mask (nonzero = implant)")
    p.add_argument("--output", required=True, help="output .npy (uint16 EDT field)")
    p.add_argument("-v", "--verbose", type=int, default=1)
    args = p.parse_args(argv)
    implant = np.load(args.implant)
    if args.verbose >= 1:
        print(f"Computing the EDT field of {implant.shape}")
    field = generate_edt_field(implant, verbose=args.verbose)
    if args.verbose >= 1:
        print(f"Writing EDT-field to {args.output}")
    np.save(args.output, field)


if __name__ == "__main__":
    main()
