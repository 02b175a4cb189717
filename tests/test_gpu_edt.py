"""GPU EDT field (maxibone_amd/csrc/mxb_edt.hip) vs the oracle
(oracle/edt.py): exact squared distances, their float32 square roots and the
0950 uint16 field, bit for bit, through the C ABI, the Python drop-in and the
device entries."""
import sys

import numpy as np
import pytest

from maxibone_amd import capi, edt, hip
from maxibone_amd.helpers import generate_cylinder_mask
from oracle import edt as E

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


SHAPES = [(6, 7, 9), (1, 1, 13), (5, 1, 8), (9, 8, 1), (17, 33, 65), (64, 64, 64), (3, 130, 70),
          (40, 20, 129), (2, 300, 5), (70, 3, 257)]


@pytest.mark.parametrize("case", range(len(SHAPES)))
def test_edt_sq_matches_oracle(gpu, case):
    rng = np.random.default_rng(900 + case)
    shape = SHAPES[case]
    density = [0.05, 0.2, 0.1, 0.3, 0.01, 0.002, 0.05, 0.02, 0.5, 0.001][case]
    data = rng.random(shape) > density
    data.flat[rng.integers(data.size)] = False
    _say("edt_sq", shape, density)
    want = E.edt_sq(data)
    got = edt.edt_sq(data)
    assert np.array_equal(got.astype(np.uint64), want)
    f = edt.edt(data)
    assert np.array_equal(f, E.edt(data))


@pytest.mark.parametrize("n", [1024, 2048, 4096])
def test_edt_long_axes(gpu, n):
    """Lines of 1024, 2048 and 4096 voxels (bands of 16, 32 and 64 per lane)
    along each axis in turn."""
    rng = np.random.default_rng(n)
    for axis in range(3):
        shape = [3, 4, 5]
        shape[axis] = n
        data = rng.random(shape) > 0.003
        data.flat[rng.integers(data.size)] = False
        _say("edt long axis", shape)
        assert np.array_equal(edt.edt_sq(data).astype(np.uint64), E.edt_sq(data))


def test_edt_no_site_and_all_sites(gpu):
    assert np.all(edt.edt_sq(np.ones((5, 6, 7), bool)) == edt.EDT_NONE)
    assert np.all(np.isinf(edt.edt(np.ones((5, 6, 7), bool))))
    assert np.all(edt.edt_sq(np.zeros((5, 6, 7), bool)) == 0)
    one = np.ones((9, 10, 11), bool)
    one[0, 0, 0] = False
    z, y, x = np.indices(one.shape)
    assert np.array_equal(edt.edt_sq(one), (z * z + y * y + x * x).astype(np.uint32))


def test_edt_ties_and_structured_sites(gpu):
    """Periodic and planar sites: many equal distances (ties in every pass)."""
    data = np.ones((24, 40, 56), bool)
    data[::7, ::5, ::3] = False
    assert np.array_equal(edt.edt_sq(data).astype(np.uint64), E.edt_sq(data))
    data = np.ones((24, 40, 56), bool)
    data[:, 13, :] = False
    data[5, :, 40] = False
    assert np.array_equal(edt.edt_sq(data).astype(np.uint64), E.edt_sq(data))


def _rod_implant(shape, rng):
    nz, ny, nx = shape
    implant = np.zeros(shape, bool)
    cy, cx, r = ny // 2, nx // 2, max(2, min(ny, nx) // 10)
    yy, xx = np.mgrid[:ny, :nx]
    implant[nz // 4:, (yy - cy) ** 2 + (xx - cx) ** 2 <= r * r] = True
    implant[rng.random(shape) < 1e-4] = True  # scattered bright specks
    return implant


@pytest.mark.parametrize("shape", [(32, 48, 48), (60, 64, 80), (128, 128, 128)])
def test_edt_field_matches_oracle(gpu, shape):
    rng = np.random.default_rng(sum(shape))
    implant = _rod_implant(shape, rng)
    cyl = generate_cylinder_mask(shape[1], shape[2])
    _say("edt field", shape)
    want = E.edt_field(implant, cyl)
    assert np.array_equal(edt.generate_edt_field(implant, cyl), want)
    assert np.array_equal(edt.generate_edt_field(implant), want)  # default cylinder mask


def test_device_entries_match_host_entries(gpu):
    rng = np.random.default_rng(1000)
    L = capi.lib()
    shape = (20, 36, 44)
    implant = _rod_implant(shape, rng)
    cyl = generate_cylinder_mask(36, 44).astype(np.uint8)
    data = np.ascontiguousarray(~implant, dtype=np.uint8)
    bufs = [hip.DeviceBuffer.from_array(a) for a in (data, implant.astype(np.uint8), cyl)]
    n = data.size
    sq = hip.DeviceBuffer(4 * n)
    fl = hip.DeviceBuffer(4 * n)
    q = hip.DeviceBuffer(2 * n)
    ws = hip.DeviceBuffer(L.mxb_edt_workspace_bytes(*shape))
    st = hip.Stream()
    capi.check(L.mxb_device_edt(bufs[0].ptr, *shape, sq.ptr, fl.ptr, ws.ptr, st.handle))
    capi.check(L.mxb_device_edt_field(bufs[1].ptr, *shape, bufs[2].ptr, q.ptr, ws.ptr, st.handle))
    st.synchronize()
    assert np.array_equal(sq.to_array(np.uint32, shape).astype(np.uint64), E.edt_sq(data))
    assert np.array_equal(fl.to_array(np.float32, shape), E.edt(data))
    assert np.array_equal(q.to_array(np.uint16, shape), E.edt_field(implant, cyl != 0))
    for b in bufs + [sq, fl, q, ws]:
        b.free()


def test_edt_field_large_host_copies(gpu):
    """A volume whose host copies go through the pinned staging chunks (136 MB
    in, 272 MB out): the host entry equals the device entry bit for bit."""
    rng = np.random.default_rng(1100)
    shape = (520, 512, 512)
    implant = _rod_implant(shape, rng)
    cyl = generate_cylinder_mask(512, 512).astype(np.uint8)
    _say("edt field staged", shape)
    host = edt.generate_edt_field(implant, cyl)
    L = capi.lib()
    dm = hip.DeviceBuffer.from_array(implant.astype(np.uint8))
    dc = hip.DeviceBuffer.from_array(cyl)
    dq = hip.DeviceBuffer(2 * implant.size)
    ws = hip.DeviceBuffer(L.mxb_edt_workspace_bytes(*shape))
    capi.check(L.mxb_device_edt_field(dm.ptr, *shape, dc.ptr, dq.ptr, ws.ptr, None))
    hip.synchronize()
    assert np.array_equal(host, dq.to_array(np.uint16, shape))
    for b in (dm, dc, dq, ws):
        b.free()
