"""GPU consumers of the diffusion field vs the oracle (oracle/fields.py):
combine_fields (0975_combine_fields.py:41-62) bit for bit, histograms
(cpu/histograms.cc) count for count, through the C ABI, the drop-in
modules and the device-resident entries."""
import sys

import numpy as np
import pytest

from maxibone_amd import capi, fields, hip, histograms
from maxibone_amd.helpers import gauss_kernel, generate_cylinder_mask
from oracle import fields as F

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fields(rng, shape, zero_frac=0.1):
    g = rng.integers(0, 65536, shape, dtype=np.uint16)
    e = rng.integers(0, 65536, shape, dtype=np.uint16)
    g[rng.random(shape) < zero_frac] = 0
    e[rng.random(shape) < zero_frac] = 0
    return g, e


@pytest.mark.parametrize("case", range(5))
def test_combine_fields_bit_exact(gpu, case):
    rng = np.random.default_rng(300 + case)
    shape = [(6, 24, 24), (5, 30, 37), (4, 16, 16), (3, 64, 64), (2, 7, 9)][case]
    g, e = _fields(rng, shape)
    cyl = [generate_cylinder_mask(24), generate_cylinder_mask(30, 37), np.ones((16, 16), bool),
           generate_cylinder_mask(64), rng.random((7, 9)) > 0.5][case]
    _say("combine", shape)
    want = F.combine_fields(g, e, cyl)
    got = fields.combine_fields(g, e, cyl)
    assert np.array_equal(got, want)


def test_combine_fields_from_a_diffusion_field(gpu):
    """The 0975 use: the uint16 Gaussian field of a diffusion and an EDT-like field."""
    from maxibone_amd.synth import WORKLOADS
    m = WORKLOADS["c1"].mask()
    gq = capi.diffusion(m, gauss_kernel(3.0).astype(np.float32), 10)[0]
    zz, yy, xx = np.meshgrid(*[np.arange(s) for s in m.shape], indexing="ij")
    edt = (np.abs(yy - 64) + np.abs(xx - 64) + (zz % 7)).astype(np.uint16) * 100
    cyl = generate_cylinder_mask(128)
    want = F.combine_fields(gq, edt, cyl)
    assert np.array_equal(fields.combine_fields(gq, edt, cyl), want)
    assert np.array_equal(fields.combine_fields(gq, edt), want)  # default cylinder mask


def test_combine_fields_rejects_zero_field(gpu):
    g = np.zeros((2, 8, 8), np.uint16)
    with pytest.raises(ValueError):
        fields.combine_fields(g, g + 1)


@pytest.mark.parametrize("case", range(6))
def test_field_histogram_counts(gpu, case):
    rng = np.random.default_rng(400 + case)
    vshape = [(20, 24, 28), (21, 16, 30), (9, 10, 11), (16, 32, 8), (7, 5, 64), (40, 64, 64)][case]
    scale = [1, 2, 1, 2, 4, 2][case]
    fshape = tuple(max(1, s // scale) for s in vshape)
    vox = rng.integers(0, 65536, vshape, dtype=np.uint16)
    vox[rng.random(vshape) < 0.2] = 0
    fld = rng.integers(0, 65536, fshape, dtype=np.uint16)
    fld[rng.random(fshape) < 0.2] = 0
    vb, fb = [(128, 64), (4096, 256), (16, 16), (100, 37), (2, 3), (4096, 2048)][case]
    vrange = [(1.0, 65535.0), (1000.0, 60000.0), (5000.0, 40000.0), (0.0, 65535.0), (1.0, 65535.0),
              (2000.0, 50000.0)][case]
    frange = [(1.0, 65535.0), (1.0, 65535.0), (1.0, 65535.0), (100.0, 30000.0), (1.0, 65535.0),
              (1.0, 65535.0)][case]
    init = rng.integers(0, 5, (fb, vb)).astype(np.uint64)
    want = F.field_histogram(vox, fld, init.copy(), vrange, frange)
    got = init.copy()
    _say("field histogram", vshape, fshape, (fb, vb))
    histograms.field_histogram(vox, fld, (0, 0, 0), got, vrange, frange, 0)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("case", range(7))
def test_axis_histograms_counts(gpu, case):
    # cases 4 and 6: more than 4096 bins (4-column blocks; case 6 with the
    # maximum 8192 bins, rows of 30 voxels and slices beyond Nz);
    # case 5: few distinct values over 300 slices (wave-combined counts, two
    # z ranges of the column kernel, slices beyond Nz skipped)
    rng = np.random.default_rng(500 + case)
    nz, ny, nx = [(12, 24, 28), (5, 40, 33), (20, 17, 17), (32, 64, 64), (10, 20, 24), (300, 6, 20),
                  (9, 11, 30)][case]
    z_off, Nz = [(0, 12), (7, 40), (3, 15), (60, 80), (0, 10), (0, 290), (2, 10)][case]
    vb = [128, 256, 64, 4096, 6000, 512, 8192][case]
    Nr = [20, 30, 5, 46, 20, 12, 9][case]  # case 2: most radii beyond Nr are skipped
    vrange = [(1.0, 65535.0), (2000.0, 50000.0), (1.0, 30000.0), (100.0, 65535.0), (1.0, 65535.0),
              (1.0, 65535.0), (1.0, 65535.0)][case]
    center = (ny // 2, nx // 2)
    if case == 5:
        vox = (rng.integers(0, 4, (nz, ny, nx)) * 1000 + 500).astype(np.uint16)
    else:
        vox = rng.integers(0, 65536, (nz, ny, nx), dtype=np.uint16)
    vox[rng.random(vox.shape) < 0.1] = 0
    bins = [rng.integers(0, 3, (n, vb)).astype(np.uint64) for n in (nx, ny, Nz, Nr)]
    want = F.axis_histograms(vox, z_off, *[b.copy() for b in bins], center, vrange)
    got = [b.copy() for b in bins]
    _say("axis histograms", (nz, ny, nx), z_off, Nz, Nr, vb)
    histograms.axis_histograms(vox, (z_off, 0, 0), *got, center, vrange, 0)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_device_entries_match_host_entries(gpu):
    rng = np.random.default_rng(600)
    L = capi.lib()
    shape = (16, 48, 48)
    g, e = _fields(rng, shape)
    cyl = generate_cylinder_mask(48).astype(np.uint8)
    bufs = [hip.DeviceBuffer.from_array(a) for a in (g, e, cyl)]
    out = hip.DeviceBuffer(g.nbytes)
    ws = hip.DeviceBuffer(L.mxb_combine_workspace_bytes())
    st = hip.Stream()
    capi.check(L.mxb_device_combine_fields(bufs[0].ptr, bufs[1].ptr, *shape, bufs[2].ptr, out.ptr, ws.ptr,
                                           st.handle))
    st.synchronize()
    assert np.array_equal(out.to_array(np.uint16, shape), F.combine_fields(g, e, cyl != 0))
    # histograms on the same device volumes
    fb, vb = 64, 512
    bins = np.zeros((fb, vb), np.uint64)
    db = hip.DeviceBuffer.from_array(bins)
    hws = hip.DeviceBuffer(L.mxb_histogram_workspace_bytes(*shape))
    capi.check(L.mxb_device_field_histogram(bufs[0].ptr, *shape, bufs[1].ptr, *shape, db.ptr, fb, vb,
                                            1.0, 65535.0, 1.0, 65535.0, hws.ptr, st.handle))
    st.synchronize()
    assert np.array_equal(db.to_array(np.uint64, (fb, vb)),
                          F.field_histogram(g, e, bins.copy(), (1.0, 65535.0), (1.0, 65535.0)))
    for b in bufs + [out, ws, db, hws]:
        b.free()
