"""Host-side helpers on the diffusion path.

``gauss_kernel`` and ``generate_cylinder_mask`` mirror the two functions of
src/lib/py/helpers.py that the diffusion caller uses (hlp:376-409 and
hlp:411-434): same names, arguments and results.
"""
import numpy as np


def gauss_kernel(sigma):
    """1D Gaussian taps in float64, as ``scipy.ndimage.gaussian_filter1d`` builds them.

    radius = round(4*sigma) with Python's round-half-to-even (hlp:392); taps
    exp(-x^2 / (2 sigma^2)) for x in [-radius, radius], normalised by their FP64
    sum (hlp:404-407).  The drop-in entry point casts them to float32, as the
    reference's ``forcecast`` argument does (datatypes.hh:56-57).
    """
    sigma = float(sigma)
    if not sigma > 0:
        raise ValueError(f"sigma must be positive, got {sigma}")
    radius = round(4.0 * sigma)
    # hlp:402-404's order, so the FP64 taps match it bit for bit: the squared
    # offsets stay integers and are scaled once by the FP64 factor -0.5/sigma^2
    offsets = np.arange(-radius, radius + 1)
    phi = np.exp((-0.5 / (sigma * sigma)) * (offsets * offsets))
    return phi / phi.sum()


def generate_cylinder_mask(ny, nx=None):
    """2D boolean disc spanning the (ny, nx) plane (hlp:411-434)."""
    if nx is None:
        nx = ny
    ys = np.linspace(-1, 1, ny)
    xs = np.linspace(-1, 1, nx)
    return (xs[np.newaxis, :] ** 2 + ys[:, np.newaxis] ** 2) < 1
