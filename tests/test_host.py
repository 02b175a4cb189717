"""Host-side code around the diffusion path: raw slices and the field driver.

The driver (maxibone_amd.gauss_field, the counterpart of
src/processing_steps/0900_generate_gauss_c.py) is exercised on the CPU with
an oracle-backed stand-in for the drop-in `diffusion` (both overloads), and
on the GPU with the real module.
"""
import os

import numpy as np
import pytest

import oracle
from maxibone_amd import gauss_field
from maxibone_amd.helpers import gauss_kernel, generate_cylinder_mask
from maxibone_amd.io import load_slice, write_slice
from maxibone_amd.synth import sphere_cylinder_mask


def test_raw_slices_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    vol = rng.integers(0, 65535, (7, 5, 6)).astype(np.uint16)
    path = str(tmp_path / "v.uint16")
    write_slice(vol[3:], path, (3, 0, 0), vol.shape)  # out of order on purpose
    write_slice(vol[:3], path, (0, 0, 0), vol.shape)
    assert os.path.getsize(path) == vol.nbytes
    got = np.empty_like(vol)
    load_slice(got, path, (0, 0, 0), vol.shape)
    assert np.array_equal(got, vol)
    part = np.empty((2, 5, 6), np.uint16)
    load_slice(part, path, (4, 0, 0), vol.shape)
    assert np.array_equal(part, vol[4:6])
    with pytest.raises(RuntimeError):
        load_slice(np.empty((5, 5, 6), np.uint16), path, (4, 0, 0), vol.shape)


def _oracle_diffusion(*args):
    """Both overloads of the drop-in `diffusion`, computed by the oracle."""
    if isinstance(args[0], str):
        input_file, kernel, output_file, total, glob, reps, *_ = args
        m = np.fromfile(input_file, np.uint8, int(np.prod(total))).reshape(total)
        u = oracle.diffusion_chunked_f32(m, np.asarray(kernel, np.float32), reps, glob[0])
        oracle.quantise_u16(u).tofile(output_file)
    else:
        mask, kernel, out, reps, *_ = args
        _, q = oracle.diffusion_f32(mask, np.asarray(kernel, np.float32), reps)
        out[:] = q


def test_field_driver_paths_agree(tmp_path):
    m = sphere_cylinder_mask((30, 24, 24), 5, 6, 2, 4, 3)
    a = gauss_field.generate_gauss_field(m, voxel_size_um=2.0, sigma_um=4.0, repetitions=3,
                                         on_disk=False, diffusion_fn=_oracle_diffusion)
    b = gauss_field.generate_gauss_field(m, voxel_size_um=2.0, sigma_um=4.0, repetitions=3,
                                         on_disk=True, workdir=str(tmp_path),
                                         diffusion_fn=_oracle_diffusion)
    assert np.array_equal(a, b)
    _, q = oracle.diffusion_f32(m, gauss_kernel(2.0).astype(np.float32), 3)
    assert np.array_equal(a, q * generate_cylinder_mask(24)[None])


def test_field_driver_memory_rule():
    assert gauss_field.host_footprint((10, 20, 30)) == 6000 * 11  # 0900:69-73
    assert gauss_field.chunk_layers((4096, 512, 512)) == 1024     # 1 GiB of float32 layers
    assert gauss_field.chunk_layers((100, 512, 512)) == 100
    assert gauss_field.chunk_layers((100, 512, 512), sample_scale=8) == 51


@pytest.mark.gpu
def test_field_driver_on_gpu(gpu, tmp_path):
    m = sphere_cylinder_mask((40, 32, 32), 6, 8, 2, 5, 4)
    a = gauss_field.generate_gauss_field(m, 1.5, 4.5, 4, on_disk=False)
    b = gauss_field.generate_gauss_field(m, 1.5, 4.5, 4, on_disk=True, workdir=str(tmp_path))
    assert np.array_equal(a, b)
    want = gauss_field.generate_gauss_field(m, 1.5, 4.5, 4, on_disk=False,
                                            diffusion_fn=_oracle_diffusion)
    assert np.abs(a.astype(np.int32) - want.astype(np.int32)).max() <= 1
    stats = gauss_field.verify_against_ndimage(m, a, 1.5, 4.5, 4)
    assert stats["max_abs"] <= 1
