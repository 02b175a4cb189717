"""Error of each precision form on crops of the workloads' own masks.

Each case runs a central crop of the workload's mask (centred on the cylinder
surface, so it holds implant, the diffusion front and far field) with the
workload's sigma and repetition count; the oracle at full size would take
hours on the host.  Bars:
* fp32 (MFMA) and fp32_split: max-abs <= 1e-5 and uint16 within 1 LSB of the
  FP32 oracle (the reference CPU arithmetic; src/test/test_diffusion.py:92-98);
* bf16, fp16: max-abs <= 2e-3 of the FP64 oracle (the FP32 taps promoted, so
  the numbers measure storage and arithmetic precision only).
Max-abs and RMS of every form against FP64 are printed (`pytest -s`) for the
record.
"""
import sys

import numpy as np
import pytest

import oracle
from maxibone_amd.synth import WORKLOADS

pytestmark = pytest.mark.gpu
TOL = 2e-3
F32_TOL = 1e-5

CASES = [  # (workload, crop edge, repetitions)
    ("c1", 128, 10),
    ("c2", 112, 40),
    ("c4", 96, 100),
]


def _crop(mask, c, cyl_r):
    """A c^3 window centred on the cylinder surface (x = nx/2 + r)."""
    nz, ny, nx = mask.shape
    oz, oy = (nz - c) // 2, (ny - c) // 2
    ox = min(max(nx // 2 + cyl_r - c // 2, 0), nx - c)
    return np.ascontiguousarray(mask[oz:oz + c, oy:oy + c, ox:ox + c])


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,edge,reps", CASES)
def test_precision_error_report(gpu, workload_mask, name, edge, reps):
    w = WORKLOADS[name]
    m = _crop(workload_mask(name), edge, w.cyl_r)
    k32 = w.kernel().astype(np.float32)
    u64 = oracle.diffusion_f64(m, k32.astype(np.float64), reps)
    u32, q32 = oracle.diffusion_f32(m, k32, reps)
    for prec in ("fp32", "fp32_split", "fp16", "bf16"):
        q, u = gpu.diffusion(m, k32, reps, precision=prec, u16=True, f32=True)
        d = u.astype(np.float64) - u64
        mx, rms = float(np.abs(d).max()), float(np.sqrt(np.mean(d * d)))
        _say(f"error-report {name} crop={edge}^3 sigma={w.sigma} reps={reps} {prec}: "
             f"max_abs={mx:.3e} rms={rms:.3e} (vs FP64)")
        if prec.startswith("fp32"):
            err = float(np.abs(u - u32).max())
            dq = int(np.abs(q.astype(np.int32) - q32.astype(np.int32)).max())
            _say(f"error-report {name} {prec} vs the FP32 oracle: max_abs={err:.3e}, uint16 {dq} LSB")
            assert err <= F32_TOL and dq <= 1
        else:
            assert mx <= TOL
