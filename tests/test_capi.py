"""The C-ABI library and the drop-in module, without a GPU.

The library must load, export every function include/maxibone_diffusion.h
declares, and reject invalid arguments before touching the device; the
pybind module must expose the reference's two overloads with the reference's
argument names (src/pybind/diffusion-pybind.cc:84-85).
"""
import ctypes
import inspect
import os

import numpy as np
import pytest

from maxibone_amd import capi
from maxibone_amd import synth


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = capi.declared_functions()
    assert len(names) >= 13
    for n in names:
        assert hasattr(L, n), n
    assert L.mxb_version().startswith(b"maxibone_amd")


def test_state_bytes():
    L = capi.lib()
    # 120 voxels rounded to 128, plus the 2048-element gap before the split
    # form's low plane, in FP32 bytes; then 4 * 5 rows of one bit word
    gap = 2048 * 4
    assert L.mxb_state_bytes(capi.PRECISION["fp32"], 4, 5, 6) == 512 + gap + 80
    assert L.mxb_state_bytes(capi.PRECISION["bf16"], 4, 5, 6) == 512 + gap + 80  # FP32-sized: the last repetition
    assert L.mxb_state_bytes(capi.PRECISION["fp32"], 2, 3, 33) == ((2 * 3 * 33 + 127) // 128 * 128 * 4 + gap
                                                                    + 2 * 3 * 2 * 4)
    assert L.mxb_state_bytes(1024, 4, 5, 6) == -1   # unknown precision
    assert L.mxb_state_bytes(capi.PRECISION["fp32"], 4, -5, 6) == -1


def test_drop_in_module_signatures():
    from maxibone_amd.diffusion import diffusion
    doc = diffusion.__doc__
    assert "np_voxels" in doc and "np_kernel" in doc and "np_output" in doc
    assert "repititions" in doc and "verbose" in doc  # the reference's spelling
    assert "input_file" in doc and "total_shape" in doc and "global_shape" in doc


def _call(*args, **kw):
    from maxibone_amd.diffusion import diffusion
    return diffusion(*args, **kw)


def test_drop_in_rejects_bad_arguments_before_device_work():
    m = np.zeros((4, 5, 6), np.uint8)
    k = np.ones(3, np.float32) / 3
    with pytest.raises(ValueError):
        _call(m, np.ones(4, np.float32), np.zeros(m.shape, np.uint16))  # even kernel
    with pytest.raises(ValueError):
        _call(np.zeros((4, 5), np.uint8), k, np.zeros((4, 5), np.uint16))  # 2-D
    with pytest.raises(ValueError):
        _call(m, k, np.zeros((4, 5, 7), np.uint16))  # output shape
    with pytest.raises(TypeError):
        _call(m, k, np.zeros(m.shape, np.uint8))  # output dtype: noconvert
    with pytest.raises(TypeError):
        _call(m, k, np.zeros((6, 5, 4), np.uint16).T)  # non-contiguous output
    with pytest.raises(ValueError):
        _call(m, k, np.zeros(m.shape, np.uint16), -1)  # negative repetitions
    with pytest.raises(ValueError):
        _call(m, np.ones((3, 1), np.float32), np.zeros(m.shape, np.uint16))


def test_on_disk_rejects_bad_arguments(tmp_path):
    k = np.ones(3, np.float32) / 3
    missing = str(tmp_path / "missing.uint8")
    with pytest.raises(RuntimeError):
        _call(missing, k, str(tmp_path / "o.uint16"), (4, 5, 6), (2, 5, 6), 1)
    short = tmp_path / "short.uint8"
    np.zeros(10, np.uint8).tofile(short)
    with pytest.raises(RuntimeError):
        _call(str(short), k, str(tmp_path / "o.uint16"), (4, 5, 6), (2, 5, 6), 1)
    with pytest.raises(ValueError):
        _call(str(short), k, str(tmp_path / "o.uint16"), (4, 5, 6), (2, 4, 6), 1)  # y split
    with pytest.raises(ValueError):
        _call(str(short), np.ones(2, np.float32), str(tmp_path / "o.uint16"), (4, 5, 6), (2, 5, 6))


def test_capi_validation_codes():
    L = capi.lib()
    m = np.zeros(8, np.uint8)
    k = np.ones(3, np.float32)
    q = np.zeros(8, np.uint16)
    assert L.mxb_diffusion_u8_u16(m.ctypes.data, 2, 2, 2, k.ctypes.data, 2, 1, q.ctypes.data, 0) == capi.EINVAL
    assert L.mxb_diffusion_u8_u16(m.ctypes.data, 2, 2, 2, k.ctypes.data, 3, -1, q.ctypes.data, 0) == capi.EINVAL
    assert L.mxb_diffusion_u8_u16(m.ctypes.data, 2, 2, 2, None, 3, 1, q.ctypes.data, 0) == capi.EINVAL
    assert L.mxb_diffusion_ex(m.ctypes.data, 2, 2, 2, k.ctypes.data, 3, 1, 99, q.ctypes.data, None, 0) == capi.EINVAL
    assert L.mxb_diffusion_ex(m.ctypes.data, 2, 2, 2, k.ctypes.data, 3, 1, 0, None, None, 0) == capi.EINVAL
    assert L.mxb_diffusion_u8_u16(m.ctypes.data, 2, 2, 2, k.ctypes.data, 513, 1, q.ctypes.data, 0) == capi.EINVAL
    assert L.mxb_diffusion_chunked(m.ctypes.data, 2, 2, 2, k.ctypes.data, 3, 1, 0, 1, 0, q.ctypes.data, None, 0) == capi.EINVAL
    assert b"chunk_z" in L.mxb_last_error()
    # empty volumes are a no-op
    assert L.mxb_diffusion_u8_u16(m.ctypes.data, 0, 2, 2, k.ctypes.data, 3, 1, q.ctypes.data, 0) == capi.OK


def test_synthetic_masks_are_deterministic_binary():
    a = synth.sphere_cylinder_mask((40, 30, 50), 6, 12, 2, 5, 9)
    b = synth.sphere_cylinder_mask((40, 30, 50), 6, 12, 2, 5, 9)
    assert np.array_equal(a, b) and set(np.unique(a)) <= {0, 1}
    yy, xx = np.mgrid[:30, :50]
    disc = (yy - 15) ** 2 + (xx - 25) ** 2 <= 36
    assert np.all(a[:, disc] == 1)
    kat = synth.kat_mask(16)
    assert kat.dtype == bool and kat.shape == (16, 16, 16)


def test_batched_volume_recipe():
    vols = synth.batched_volumes(n_volumes=5, seed=3)
    assert len(vols) == 5
    for v in vols:
        assert all(64 <= s <= 384 for s in v.mask.shape)
        assert 2 <= v.sigma <= 10 and v.radius == round(4 * v.sigma)


def test_edt_argument_checks_before_the_device():
    """The EDT entries reject oversized axes and missing outputs before any
    device work (include/maxibone_edt.h: at most 4096 voxels per axis)."""
    from maxibone_amd import edt as E
    L = capi.lib()
    assert L.mxb_edt_workspace_bytes(4, 5, 6) == 2 * 512 + 256  # 480 B volumes rounded to 512
    assert L.mxb_edt_workspace_bytes(-1, 5, 6) == -1
    big = np.zeros(1, np.uint8)
    out = np.zeros(1, np.float32)
    assert L.mxb_edt(big.ctypes.data, E.MAX_AXIS + 1, 1, 1, None, out.ctypes.data, 0) == capi.EINVAL
    assert L.mxb_edt(big.ctypes.data, 1, 1, 1, None, None, 0) == capi.EINVAL
    assert L.mxb_edt_field(big.ctypes.data, 1, E.MAX_AXIS + 1, 1, None, None, 0) == capi.EINVAL
    with pytest.raises(ValueError):
        E.edt(np.zeros((4, 4), bool))


def test_header_states_the_noalloc_scratch_size():
    import re
    text = open(capi.HEADERS[0]).read()
    m = re.search(r"#define MXB_NOALLOC_SCRATCH_BYTES (\d+)", text)
    assert m and int(m.group(1)) == capi.NOALLOC_SCRATCH_BYTES
