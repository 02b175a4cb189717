"""Raw, headerless C-order volume slices: the file format of the chunked path.

Mirrors the reference's `load_slice` / `write_slice`
(src/pybind/io-pybind.cc:24-60, src/lib/cpp/cpu_seq/io.cc:12-40): a file is a
flat array of one element type; a slice at `offset` (z, y, x) of a volume of
`shape` starts at element oz*Ny*Nx + oy*Nx + ox and covers the whole of the
destination (or source) array.  Implemented with numpy memory maps.
"""
import os

import numpy as np


def _flat_offset(offset, shape):
    oz, oy, ox = (int(v) for v in offset)
    _, ny, nx = (int(v) for v in shape)
    return oz * ny * nx + oy * nx + ox


def load_slice(data, filename, offset, shape):
    """Fill `data` (any shape, C-contiguous) from `filename` at the slice offset."""
    if not data.flags.c_contiguous:
        raise ValueError("data must be C-contiguous")
    start = _flat_offset(offset, shape)
    itemsize = data.dtype.itemsize
    size = os.path.getsize(filename)
    if (start + data.size) * itemsize > size:
        raise RuntimeError(f"{filename}: holds {size} bytes, slice needs {(start + data.size) * itemsize}")
    src = np.memmap(filename, dtype=data.dtype, mode="r", offset=start * itemsize, shape=(data.size,))
    data.reshape(-1)[:] = src
    del src


def write_slice(data, filename, offset, shape):
    """Write `data` into `filename` at the slice offset, creating/growing the file."""
    data = np.ascontiguousarray(data)
    start = _flat_offset(offset, shape)
    itemsize = data.dtype.itemsize
    end = (start + data.size) * itemsize
    mode = "r+b" if os.path.exists(filename) else "w+b"
    with open(filename, mode) as f:
        f.seek(0, os.SEEK_END)
        if f.tell() < end:
            f.truncate(end)
        f.seek(start * itemsize)
        f.write(data.tobytes(order="C"))
