"""Host driver for the implant "glow" field: the counterpart of
src/processing_steps/0900_generate_gauss_c.py.

Steps (reference lines in brackets):
1. sigma in voxels = sigma in micrometres / voxel size             [0900:56]
2. FP64 Gaussian taps (cast to float32 by the drop-in module)      [0900:66]
3. choose in-memory vs file-backed by the host-memory footprint
   N * (1 + 2*4 + 2) bytes against the available memory             [0900:68-75]
4. file-backed: write the mask as raw uint8, run the on-disk overload
   with chunks of floor(1 GiB / 4 / (ny*nx)) layers, read the uint16
   result back                                                      [0900:76-104]
   in-memory: diffusion(mask, kernel, result, repetitions, verbose) [0900:105-115]
5. multiply by the 2D cylinder mask and save as .npy                [0900:117-124]
6. optional --verify: repeated scipy.ndimage.gaussian_filter with
   illumination, uint16 difference statistics                      [0900:126-159]

Inputs are .npy masks (or raw uint8 with --shape); the reference reads an
HDF5 file, and h5py is not a dependency of this package.
"""
import argparse
import os
import tempfile
import time

import numpy as np

from .helpers import gauss_kernel, generate_cylinder_mask
from .io import load_slice, write_slice

GIGABYTE = 1024 ** 3


def host_footprint(shape, mask_itemsize=1):
    """Bytes the in-memory path needs on the host (0900:69-73)."""
    n = int(np.prod(shape, dtype=np.int64))
    return n * mask_itemsize + 2 * n * 4 + n * 2


def chunk_layers(shape, sample_scale=1):
    """Layers per chunk of the file-backed path (0900:83-89)."""
    nz, ny, nx = shape
    if sample_scale == 8:
        return nz // 2 + 1
    layers = int(np.floor((GIGABYTE / 4) / (ny * nx)))
    return max(1, min(layers, nz))


def generate_gauss_field(implant_mask, voxel_size_um, sigma_um=10.0, repetitions=42, verbose=0,
                         on_disk=None, workdir=None, sample_scale=1, diffusion_fn=None):
    """Return the uint16 diffusion field times the cylinder mask.

    `on_disk` None chooses by available host memory (as 0900 does); True or
    False forces the path.  `diffusion_fn` defaults to the drop-in module's
    `diffusion` (both overloads).
    """
    if diffusion_fn is None:
        from .diffusion import diffusion as diffusion_fn
    mask = np.ascontiguousarray(implant_mask)
    if mask.ndim != 3:
        raise ValueError("implant_mask must be 3-dimensional")
    nz, ny, nx = mask.shape
    sigma_voxels = sigma_um / voxel_size_um
    kernel = gauss_kernel(sigma_voxels)
    if on_disk is None:
        import psutil
        on_disk = host_footprint(mask.shape, mask.dtype.itemsize) > psutil.virtual_memory().available
    t0 = time.perf_counter()
    if on_disk:
        with tempfile.TemporaryDirectory(dir=workdir) as td:
            input_path = os.path.join(td, "implant_solid.uint8")
            output_path = os.path.join(td, "gauss.uint16")
            write_slice(mask.astype(np.uint8), input_path, (0, 0, 0), mask.shape)
            layers = chunk_layers(mask.shape, sample_scale)
            if verbose >= 1:
                print(f"Using {layers} layers of {ny}x{nx} slices")
            diffusion_fn(input_path, kernel, output_path, (nz, ny, nx), (layers, ny, nx), repetitions,
                         verbose)
            result = np.empty((nz, ny, nx), dtype=np.uint16)
            load_slice(result, output_path, (0, 0, 0), (nz, ny, nx))
    else:
        result = np.zeros(mask.shape, dtype=np.uint16)
        diffusion_fn(mask, kernel, result, repetitions, verbose)
    if verbose >= 1:
        print(f"Diffusion took {time.perf_counter() - t0:.02f} seconds")
    cylinder = generate_cylinder_mask(ny, nx)
    return result * cylinder[np.newaxis, :, :]


def verify_against_ndimage(implant_mask, result, voxel_size_um, sigma_um, repetitions,
                           cylinder=True):
    """0900:126-159: scipy gaussian_filter with illumination; difference statistics."""
    from scipy import ndimage as ndi
    sigma_voxels = sigma_um / voxel_size_um
    a = implant_mask > 0
    control = a.astype(np.float32)
    for _ in range(repetitions):
        control[:] = ndi.gaussian_filter(control, sigma_voxels, mode="constant", cval=0)
        control[a] = 1
    control = np.floor(control * np.iinfo(np.uint16).max).astype(np.uint16)
    if cylinder:
        control = control * generate_cylinder_mask(*implant_mask.shape[1:])[np.newaxis]
    diff = result.astype(np.int32) - control.astype(np.int32)
    d = np.abs(diff)
    return {"total": int(d.sum()), "max_abs": int(d.max()), "min": int(diff.min()),
            "max": int(diff.max()), "mean_abs": float(d.mean())}


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("mask", help=".npy implant mask, or raw uint8 with --shape")
    p.add_argument("sigma", type=float, nargs="?", default=10.0, help="sigma in micrometres (default 10)")
    p.add_argument("repetitions", type=int, nargs="?", default=42, help="repetitions (default 42)")
    p.add_argument("--voxel-size", type=float, required=True, help="voxel size in micrometres")
    p.add_argument("--shape", type=int, nargs=3, help="(nz ny nx) of a raw uint8 mask")
    p.add_argument("--output", required=True, help="output .npy")
    p.add_argument("--on-disk", choices=["auto", "yes", "no"], default="auto")
    p.add_argument("--verify", action="store_true")
    p.add_argument("-v", "--verbose", type=int, default=1)
    args = p.parse_args(argv)
    if args.mask.endswith(".npy"):
        mask = np.load(args.mask)
    else:
        if not args.shape:
            p.error("--shape is required for raw masks")
        mask = np.fromfile(args.mask, dtype=np.uint8).reshape(args.shape)
    on_disk = {"auto": None, "yes": True, "no": False}[args.on_disk]
    result = generate_gauss_field(mask, args.voxel_size, args.sigma, args.repetitions, args.verbose,
                                  on_disk=on_disk)
    np.save(args.output, result)
    if args.verify:
        stats = verify_against_ndimage(mask, result, args.voxel_size, args.sigma, args.repetitions)
        print(stats)


if __name__ == "__main__":
    main()
