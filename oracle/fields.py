"""Oracle for the consumers of the diffusion field (test infrastructure only).

A numpy restatement of the reference's downstream steps, used only by the
tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg as the
checker:

* ``combine_fields``: src/processing_steps/0975_combine_fields.py:41-62
  (normalise the Gaussian and EDT fields by their maxima, subtract, shift to
  zero, rescale to 1, cylinder mask, ``to_int`` to uint16, cylinder mask).
* ``normalized_convert``: src/lib/cpp/include/general.hh:96-118 (the C++
  ``to_int``): out = (U)(((float)in - vmin) / (vmax - vmin) * max(U)) in
  float32, truncated toward zero; vmin, vmax the array's own min and max.
* ``field_histogram``: src/lib/cpp/cpu/histograms.cc:231-306 (2D histogram of
  voxel value bin x field value bin, field sampled at floor(Z*dz) etc.).
* ``axis_histograms``: src/lib/cpp/cpu/histograms.cc:14-78 (per-x, -y, -z
  and radius histograms of voxel value bins).

Pinned (tests/test_fields_oracle.py) bit for bit against the reference's own
CPU modules compiled from their sources into oracle/_ref (histograms,
general), and by closed forms.  The reference's OpenACC GPU axis histogram
rounds the voxel bin (src/lib/cpp/gpu/histograms.cc:94) where the CPU and
sequential forms floor it; this restatement follows the CPU forms.  Bins the
reference would index out of range (field values outside [f_min, f_max], radii
beyond Nr) are skipped here (the reference writes out of bounds).
"""
import numpy as np


def normalized_convert(x, dtype=np.uint16):
    """general.hh:96-118 with min_max of the input (general.hh:75-83, :113-118).

    min_max starts vmax at std::numeric_limits<float>::min() (general.hh:78),
    the smallest positive normal float, so for float input the maximum used
    is max(x.max(), 1.1754944e-38): an all-negative input is scaled as if its
    maximum were (almost) zero."""
    x = np.asarray(x, np.float32)
    vmin = np.float32(x.min())
    vmax = np.float32(max(x.max(), np.finfo(np.float32).tiny))
    scale = np.float32(np.iinfo(dtype).max)
    with np.errstate(invalid="ignore", divide="ignore"):
        y = (x - vmin) / (vmax - vmin) * scale
    return y.astype(dtype)


def combine_fields(gauss, edt, cylinder):
    """0975:41-62 on uint16 fields; `cylinder` is the 2D (ny, nx) boolean mask
    (0975:39 builds it with generate_cylinder_mask(nx))."""
    g = np.asarray(gauss, np.uint16)
    e = np.asarray(edt, np.uint16)
    cyl = np.asarray(cylinder, bool)[np.newaxis]
    combined = g.astype(np.float32) / np.float32(g.max())
    combined -= e.astype(np.float32) / np.float32(e.max())
    combined += np.abs(combined.min())
    combined /= combined.max()
    combined *= cyl
    out = normalized_convert(combined, np.uint16)
    out *= cyl
    return out


def _bin_index(values, lo, hi, nbins):
    """floor((nbins-1) * ((v - lo) / (hi - lo))) in float64 (histograms.cc:62, :279, :292)."""
    v = values.astype(np.float64)
    return np.floor(np.float64(nbins - 1) * ((v - lo) / (hi - lo))).astype(np.int64)


def field_histogram(voxels, field, bins, vrange, frange):
    """histograms.cc:231-306 (accumulates into `bins`, shape (field_bins, voxel_bins))."""
    voxels = np.asarray(voxels, np.uint16)
    field = np.asarray(field, np.uint16)
    nZ, nY, nX = voxels.shape
    nz, ny, nx = field.shape
    fb, vb = bins.shape
    vmin, vmax = vrange
    fmin, fmax = frange
    dz, dy, dx = nz / nZ, ny / nY, nx / nX
    v = voxels.astype(np.float64)
    v = np.where((v >= vmin) & (v <= vmax), v, 0.0)
    vi = np.floor(np.float64(vb - 1) * ((v - vmin) / (vmax - vmin))).astype(np.int64)
    z = np.minimum(np.floor(np.arange(nZ, dtype=np.float64) * dz).astype(np.int64), nz - 1)
    y = np.floor(np.arange(nY, dtype=np.float64) * dy).astype(np.int64)
    x = np.floor(np.arange(nX, dtype=np.float64) * dx).astype(np.int64)
    f = field[z[:, None, None], y[None, :, None], x[None, None, :]]
    fi = _bin_index(f, fmin, fmax, fb)
    keep = (v != 0) & (f > 0) & (fi >= 0) & (fi < fb) & (vi >= 0) & (vi < vb)
    flat = fi[keep] * vb + vi[keep]
    bins += np.bincount(flat, minlength=fb * vb).reshape(fb, vb).astype(np.uint64)
    return bins


def axis_histograms(voxels, offset_z, x_bins, y_bins, z_bins, r_bins, center, vrange):
    """histograms.cc:14-78 (accumulates; voxels is a z-chunk starting at
    global layer `offset_z`; bins shapes (Nx, vb), (Ny, vb), (Nz, vb), (Nr, vb))."""
    voxels = np.asarray(voxels, np.uint16)
    nz, ny, nx = voxels.shape
    vb = x_bins.shape[1]
    Nr = r_bins.shape[0]
    cy, cx = center
    vmin, vmax = vrange
    v = voxels.astype(np.float64)
    v = np.where((v >= vmin) & (v <= vmax), v, 0.0)
    vi = np.floor(np.float64(vb - 1) * ((v - vmin) / (vmax - vmin))).astype(np.int64)
    keep = (v != 0) & (vi < vb)
    # layers beyond the global extent are skipped (z_end = min(z_start + nz, Nz), :40)
    keep &= (np.arange(nz)[:, None, None] + int(offset_z)) < z_bins.shape[0]
    zz, yy, xx = np.nonzero(keep)
    idx = vi[keep]
    ddy = (yy.astype(np.int64) - int(cy)) ** 2
    ddx = (xx.astype(np.int64) - int(cx)) ** 2
    r = np.floor(np.sqrt((ddx + ddy).astype(np.float64))).astype(np.int64)
    for bins, coord in ((x_bins, xx), (y_bins, yy), (z_bins, zz + int(offset_z))):
        np.add.at(bins, (coord, idx), np.uint64(1))
    ok = r < Nr
    np.add.at(r_bins, (r[ok], idx[ok]), np.uint64(1))
    return x_bins, y_bins, z_bins, r_bins
