"""The EDT-field oracle (oracle/edt.py, oracle/edt_oracle.c) pinned against
brute force, scipy.ndimage's exact EDT and closed forms.  0950_generate_edt.py
calls the external `edt` package, which is not importable here; these checks
fix the mathematics its result is defined by (squared Euclidean distance to
the nearest zero voxel, the volume edge not a boundary)."""
import numpy as np
import pytest
from scipy import ndimage

from maxibone_amd.helpers import generate_cylinder_mask
from oracle import edt as E
from oracle import fields as F


def brute_sq(data):
    d = np.asarray(data) != 0
    sites = np.argwhere(~d)
    out = np.full(d.shape, E.INF, np.uint64)
    if len(sites) == 0:
        return out
    grid = np.indices(d.shape).reshape(3, -1).T
    best = np.full(len(grid), np.iinfo(np.int64).max, np.int64)
    for s in sites:
        best = np.minimum(best, ((grid - s) ** 2).sum(axis=1))
    return best.reshape(d.shape).astype(np.uint64)


def scipy_sq(data):
    d = np.asarray(data) != 0
    idx = ndimage.distance_transform_edt(d, return_distances=False, return_indices=True)
    grid = np.indices(d.shape)
    return ((grid - idx).astype(np.int64) ** 2).sum(axis=0).astype(np.uint64)


@pytest.mark.parametrize("case", range(10))
def test_edt_sq_matches_brute_force(case):
    rng = np.random.default_rng(700 + case)
    shape = [(6, 7, 9), (1, 1, 13), (5, 1, 8), (9, 8, 1), (4, 6, 5), (7, 7, 7), (3, 10, 4), (2, 2, 2),
             (8, 5, 6), (1, 9, 9)][case]
    density = [0.05, 0.2, 0.1, 0.3, 0.01, 0.5, 0.02, 0.5, 0.9, 0.03][case]
    data = rng.random(shape) > density  # zeros (sites) with probability `density`
    if case == 4:
        data[:] = True
        data[2, 3, 1] = False  # a single site
    assert np.array_equal(E.edt_sq(data), brute_sq(data))


def test_edt_sq_no_site_and_all_sites():
    assert np.all(E.edt_sq(np.ones((3, 4, 5), bool)) == E.INF)
    assert np.all(E.edt_sq(np.zeros((3, 4, 5), bool)) == 0)
    f = E.edt(np.ones((2, 2, 2), bool))
    assert f.dtype == np.float32 and np.all(np.isinf(f))


@pytest.mark.parametrize("case", range(4))
def test_edt_sq_matches_scipy(case):
    rng = np.random.default_rng(800 + case)
    shape = [(30, 40, 50), (64, 33, 17), (20, 64, 64), (48, 48, 48)][case]
    density = [1e-3, 1e-2, 5e-4, 0.05][case]
    data = rng.random(shape) > density
    data[rng.integers(shape[0]), rng.integers(shape[1]), rng.integers(shape[2])] = False
    assert np.array_equal(E.edt_sq(data), scipy_sq(data))


def test_edt_sq_closed_form_single_site():
    shape = (11, 13, 17)
    data = np.ones(shape, bool)
    data[4, 9, 2] = False
    z, y, x = np.indices(shape)
    want = (z - 4) ** 2 + (y - 9) ** 2 + (x - 2) ** 2
    assert np.array_equal(E.edt_sq(data), want.astype(np.uint64))


def test_edt_sq_closed_form_plane_of_sites():
    """Sites on the plane x = 5: the distance is |x - 5| everywhere."""
    data = np.ones((6, 7, 12), bool)
    data[:, :, 5] = False
    want = np.broadcast_to((np.arange(12) - 5) ** 2, data.shape)
    assert np.array_equal(E.edt_sq(data), want.astype(np.uint64))


def test_edt_float_is_float32_sqrt():
    rng = np.random.default_rng(9)
    data = rng.random((10, 12, 14)) > 0.02
    d2 = E.edt_sq(data)
    f = E.edt(data)
    assert f.dtype == np.float32
    assert np.array_equal(f, np.sqrt(d2.astype(np.float32)))


def test_edt_field_properties():
    """0950:41-66 on a rod implant: zero outside the cylinder, 0 on the
    implant, 65535 at the largest distance inside the cylinder, and
    non-decreasing in the distance."""
    n, nz = 40, 6
    implant = np.zeros((nz, n, n), bool)
    implant[:, 18:22, 18:22] = True
    cyl = generate_cylinder_mask(n)
    q = E.edt_field(implant, cyl)
    assert q.dtype == np.uint16
    assert np.all(q[:, ~cyl] == 0)
    assert np.all(q[implant] == 0)
    d = E.edt(~implant)
    inside = np.broadcast_to(cyl, q.shape)
    assert q[inside].max() == 65535
    assert q[inside][np.argmax(d[inside])] == 65535
    order = np.argsort(d[inside], kind="stable")
    assert np.all(np.diff(q[inside][order].astype(np.int64)) >= 0)
    # the conversion is the C++ to_int of the masked float field
    f = d * cyl
    assert np.array_equal(q, F.normalized_convert(f) * cyl)
