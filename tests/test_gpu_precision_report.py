"""Error of the reduced-precision forms against the FP64 oracle, per workload.

Each case runs a central crop of the workload's own mask with the workload's
sigma and repetition count (the FP64 oracle at full size would take hours on
the host) and checks max |u - u64| <= 2e-3; max-abs and RMS are printed
(`pytest -s`) for the record.  The plain FP32 form is reported too.
The oracle uses the float32 taps promoted to float64, so the numbers measure
storage and arithmetic precision only.
"""
import sys

import numpy as np
import pytest

import oracle
from maxibone_amd.synth import WORKLOADS

pytestmark = pytest.mark.gpu
TOL = 2e-3

CASES = [  # (workload, crop edge, repetitions)
    ("c1", 128, 10),
    ("c2", 112, 40),
    ("c4", 96, 100),
]


def _crop(mask, c, cyl_r):
    """A c^3 window centred on the cylinder surface (x = nx/2 + r), so that it
    holds implant, the diffusion front and far field."""
    nz, ny, nx = mask.shape
    oz, oy = (nz - c) // 2, (ny - c) // 2
    ox = min(max(nx // 2 + cyl_r - c // 2, 0), nx - c)
    return np.ascontiguousarray(mask[oz:oz + c, oy:oy + c, ox:ox + c])


@pytest.mark.parametrize("name,edge,reps", CASES)
def test_precision_error_report(gpu, name, edge, reps):
    w = WORKLOADS[name]
    if name == "c4":
        # the 2048x2048x1536 mask is large: a 512^3 volume of the same recipe
        # (cylinder r=175, sphere density and radii of the full mask)
        from maxibone_amd.synth import sphere_cylinder_mask
        mask = sphere_cylinder_mask((512, 512, 512), 175, 20000 * 512 ** 3 // (2048 * 2048 * 1536),
                                    5, 60, w.seed)
        m = _crop(mask, edge, 175)
    else:
        mask = w.mask()
        m = _crop(mask, edge, w.cyl_r)
    k32 = w.kernel().astype(np.float32)
    u64 = oracle.diffusion_f64(m, k32.astype(np.float64), reps)
    for prec in ("fp32", "fp16", "bf16"):
        u = gpu.diffusion(m, k32, reps, precision=prec, u16=False, f32=True)[1]
        d = u.astype(np.float64) - u64
        mx, rms = float(np.abs(d).max()), float(np.sqrt(np.mean(d * d)))
        print(f"error-report {name} crop={edge}^3 sigma={w.sigma} reps={reps} {prec}: "
              f"max_abs={mx:.3e} rms={rms:.3e}", flush=True)
        sys.stdout.flush()
        assert mx <= TOL
