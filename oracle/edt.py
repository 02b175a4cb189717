"""Oracle for the EDT field (test infrastructure only).

Restates src/processing_steps/0950_generate_edt.py:44-66 for the checker:

* ``edt_sq(data)``: exact squared Euclidean distance of every nonzero voxel
  to the nearest zero voxel (uint64; ``INF`` where the volume has no zero
  voxel), from ``edt_oracle.c``: the distance transform of the external
  ``edt`` package (seung-lab edt 2.4.0, src/requirements.txt:2) called by
  0950:49 as ``edt.edt(mask, parallel=n)`` with its defaults (anisotropy
  (1, 1, 1), black_border=False).
* ``edt(data)``: the package's float32 result, the float32 square root of the
  float32 squared distance (+inf where ``INF``).
* ``edt_field(implant_solid, cylinder)``: 0950:44-66: negate the implant mask
  (0950:41), EDT, multiply by the cylinder mask, ``to_int`` to uint16 (the
  C++ normalized_convert, general.hh:96-118, as ``fields.normalized_convert``),
  multiply by the cylinder mask again.

``edt_sq`` is pinned (tests/test_edt_oracle.py) against brute force on small
volumes and against scipy.ndimage.distance_transform_edt (an independent
exact EDT, Maurer et al.) on larger ones.  The ``edt`` package itself is not
importable here, so agreement with its float rounding for squared distances
above 2^24 (not reached below ~2400 voxels per axis) and its result for a
volume with no zero voxel are "parity unpinned".
"""
import numpy as np

from . import lib
from .fields import normalized_convert

INF = np.iinfo(np.uint64).max


def edt_sq(data):
    d = np.ascontiguousarray(np.asarray(data) != 0, dtype=np.uint8)
    out = np.empty(d.shape, np.uint64)
    if d.size:
        rc = lib().oracle_edt_sq(d, *d.shape, out)
        if rc != 0:
            raise MemoryError("oracle_edt_sq")
    return out


def edt(data):
    d2 = edt_sq(data)
    none = d2 == INF
    f = np.sqrt(np.where(none, 0, d2).astype(np.float32))
    f[none] = np.inf
    return f


def edt_field(implant_solid, cylinder):
    """0950:41-66 on a boolean implant mask and a (ny, nx) cylinder mask."""
    implant_mask = ~np.asarray(implant_solid, bool)         # 0950:41
    fedt = edt(implant_mask)                                 # 0950:49
    cyl = np.asarray(cylinder, bool)
    with np.errstate(invalid="ignore"):
        fedt *= cyl                                          # 0950:53
    fedt = normalized_convert(fedt, np.uint16)               # 0950:60
    fedt *= cyl                                              # 0950:65
    return fedt
