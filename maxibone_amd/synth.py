"""Seeded synthetic implant masks for the diffusion workloads.

Every mask is C-order (z, y, x) uint8 in {0, 1}: an optional solid cylinder
with its axis along z through (ny//2, nx//2), plus spheres with integer
centres uniform in the volume and integer radii uniform in [r_lo, r_hi],
clipped at the volume edges.  Membership uses exact integer arithmetic
(d^2 <= r^2), so any implementation rasterises identical masks.

``kat_mask`` reproduces the mask of the reference's known-answer test
(src/test/test_diffusion.py:32-34: legacy ``np.random.seed(42)``,
``rand(n, n, n) > 0.99``).
"""
from dataclasses import dataclass, field

import numpy as np

from .helpers import gauss_kernel


@dataclass(frozen=True)
class Workload:
    name: str
    shape: tuple
    cyl_r: int
    n_spheres: int
    r_lo: int
    r_hi: int
    sigma: float
    reps: int
    seed: int
    description: str = field(default="", compare=False)

    @property
    def radius(self):
        return round(4.0 * self.sigma)

    @property
    def voxels(self):
        return int(np.prod(self.shape, dtype=np.int64))

    def kernel(self):
        return gauss_kernel(self.sigma)

    def mask(self):
        return sphere_cylinder_mask(self.shape, self.cyl_r, self.n_spheres, self.r_lo,
                                    self.r_hi, self.seed)


WORKLOADS = {
    "c1": Workload("c1", (128, 128, 128), 20, 40, 3, 10, 3.0, 10, 1001,
                   "128^3, cylinder r=20 + 40 spheres r~U[3,10], sigma=3 (R=12), 10 reps"),
    "c2": Workload("c2", (1024, 1024, 1024), 300, 2000, 5, 40, 10.0, 40, 1002,
                   "1024^3, cylinder r=300 + 2000 spheres r~U[5,40], sigma=10 (R=40), 40 reps"),
    "c4": Workload("c4", (2048, 2048, 1536), 700, 20000, 5, 60, 15.0, 100, 1004,
                   "2048x2048x1536, cylinder r=700 + 20000 spheres r~U[5,60], sigma=15 (R=60), "
                   "100 reps"),
}


def sphere_cylinder_mask(shape, cyl_r, n_spheres, r_lo, r_hi, seed):
    nz, ny, nx = (int(s) for s in shape)
    rng = np.random.default_rng(seed)
    centres = np.stack([rng.integers(0, nz, n_spheres), rng.integers(0, ny, n_spheres),
                        rng.integers(0, nx, n_spheres)], axis=1) if n_spheres else np.zeros((0, 3), int)
    radii = rng.integers(r_lo, r_hi + 1, n_spheres) if n_spheres else np.zeros(0, int)
    mask = np.zeros((nz, ny, nx), np.uint8)
    if cyl_r > 0:
        cy, cx = ny // 2, nx // 2
        yy = (np.arange(ny, dtype=np.int64) - cy) ** 2
        xx = (np.arange(nx, dtype=np.int64) - cx) ** 2
        disc = (yy[:, None] + xx[None, :]) <= cyl_r * cyl_r
        mask[:] = disc[None, :, :]
    for (cz, cy, cx), r in zip(centres, radii):
        z0, z1 = max(cz - r, 0), min(cz + r + 1, nz)
        y0, y1 = max(cy - r, 0), min(cy + r + 1, ny)
        x0, x1 = max(cx - r, 0), min(cx + r + 1, nx)
        dz = (np.arange(z0, z1, dtype=np.int64) - cz) ** 2
        dy = (np.arange(y0, y1, dtype=np.int64) - cy) ** 2
        dx = (np.arange(x0, x1, dtype=np.int64) - cx) ** 2
        ball = (dz[:, None, None] + dy[None, :, None] + dx[None, None, :]) <= r * r
        mask[z0:z1, y0:y1, x0:x1] |= ball.astype(np.uint8)
    return mask


def kat_mask(n=128, seed=42, p=0.99):
    """The reference KAT mask (src/test/test_diffusion.py:32-34)."""
    state = np.random.RandomState(seed)
    return (state.rand(n, n, n) > p)


@dataclass
class BatchVolume:
    mask: np.ndarray
    sigma: float

    @property
    def radius(self):
        return round(4.0 * self.sigma)

    def kernel(self):
        return gauss_kernel(self.sigma)


def batched_volumes(n_volumes=64, dim_lo=64, dim_hi=384, s_lo=5, s_hi=50, r_lo=3, r_hi=20,
                    sigma_lo=2.0, sigma_hi=10.0, seed=1005):
    """Batched workload: uneven sub-volumes, each with its own sigma.

    Defaults: 64 volumes, each dimension U[64,384], 5-50 spheres r~U[3,20],
    sigma~U[2,10] with radius round(4 sigma).
    """
    rng = np.random.default_rng(seed)
    vols = []
    for i in range(n_volumes):
        shape = tuple(int(v) for v in rng.integers(dim_lo, dim_hi + 1, 3))
        ns = int(rng.integers(s_lo, s_hi + 1))
        sigma = float(rng.uniform(sigma_lo, sigma_hi))
        m = sphere_cylinder_mask(shape, 0, ns, r_lo, r_hi, seed * 1000 + i)
        vols.append(BatchVolume(m, sigma))
    return vols
