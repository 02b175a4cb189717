"""Write the golden fixtures for the diffusion path (run in the build container).

The expected outputs come from scipy.ndimage, the library routine the
reference's own known-answer test uses as its oracle
(src/test/test_diffusion.py:49-59: per repetition ``convolve1d`` along axes
0, 1, 2 with mode='constant', then ``buf[a] = 1``; result
``floor(buf * 65535)`` as uint16).  Nothing from the reference tree is
imported or executed here; the masks and taps are rebuilt from their recipes.

Fixtures (all small, numpy ``.npz``):

* ``kat_128.npz`` - the reference KAT: 128^3, ``np.random.seed(42)``,
  ``rand > 0.99``, sigma = 3 (25 taps), 100 repetitions.  Stored: a hash of
  the mask, the uint16 result on the central z-slice and on a stride-5
  subsample, and the full uint16 result's sha256.
* ``c1_128.npz`` - the 128^3 cylinder + spheres workload, sigma = 3,
  10 repetitions: same sampling, plus float32 values on the same samples.
* ``small.npz`` - small whole volumes (edge shapes, kernels longer than the
  volume, empty and full masks, zero repetitions), with float32 and uint16
  results.
* ``correlate.npz`` - asymmetric taps run through ``correlate1d``: the
  reference computes a correlation (cpu:50-54), which a flipped kernel fails.
* ``taps.npz`` - impulse responses of ``gaussian_filter1d`` (truncate = 4)
  for sigmas whose radius int(4 sigma + 0.5) equals round(4 sigma).

Usage: ``python tests/golden/make_golden.py``.
"""
import hashlib
import os
import sys

import numpy as np
import scipy.ndimage as ndi

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from maxibone_amd.helpers import gauss_kernel  # noqa: E402
from maxibone_amd.synth import WORKLOADS, kat_mask  # noqa: E402


def scipy_diffusion(mask, kernel32, reps, op=ndi.convolve1d):
    a = mask > 0
    buf0 = a.astype(np.float32)
    for _ in range(reps):
        buf1 = op(buf0, kernel32, axis=0, mode="constant")
        buf0 = op(buf1, kernel32, axis=1, mode="constant")
        buf1 = op(buf0, kernel32, axis=2, mode="constant")
        buf1[a] = 1
        buf0 = buf1
    return buf0, np.floor(buf0 * 65535).astype(np.uint16)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def sampled(u, q, mask):
    n = mask.shape[0]
    return dict(mask_sha=sha(np.packbits(mask.astype(bool))),
                slice_q=q[n // 2].copy(), slice_u=u[n // 2].copy(),
                sub_q=q[::5, ::5, ::5].copy(), sub_u=u[::5, ::5, ::5].copy(),
                full_q_sha=sha(q))


def main():
    # 1. the reference KAT
    m = kat_mask()
    k = gauss_kernel(3.0).astype(np.float32)
    u, q = scipy_diffusion(m, k, 100)
    np.savez_compressed(os.path.join(HERE, "kat_128.npz"), kernel=k, reps=100, sigma=3.0,
                        **sampled(u, q, m))

    # 2. the 128^3 cylinder + spheres workload
    w = WORKLOADS["c1"]
    m = w.mask()
    k = w.kernel().astype(np.float32)
    u, q = scipy_diffusion(m, k, w.reps)
    np.savez_compressed(os.path.join(HERE, "c1_128.npz"), kernel=k, reps=w.reps, sigma=w.sigma,
                        **sampled(u, q, m))

    # 3. small whole volumes
    rng = np.random.default_rng(7)
    cases = [
        ("single", np.ones((1, 1, 1), np.uint8), 1.0, 3),
        ("tiny_rand", (rng.random((5, 7, 9)) > 0.8).astype(np.uint8), 1.0, 4),
        ("kernel_longer", (rng.random((3, 4, 50)) > 0.9).astype(np.uint8), 1.5, 3),
        ("flat_y", (rng.random((17, 1, 33)) > 0.9).astype(np.uint8), 2.0, 3),
        ("empty", np.zeros((9, 10, 11), np.uint8), 1.0, 3),
        ("full", np.ones((6, 5, 4), np.uint8), 1.0, 2),
        ("mid", (rng.random((12, 20, 70)) > 0.97).astype(np.uint8), 2.5, 6),
        ("odd_x", (rng.random((13, 11, 129)) > 0.99).astype(np.uint8), 4.0, 3),
        ("zero_reps", (rng.random((8, 9, 10)) > 0.5).astype(np.uint8), 1.0, 0),
        ("impulse", np.pad(np.ones((1, 1, 1), np.uint8), 12), 3.0, 1),
        ("mask_values", (rng.integers(0, 256, (6, 7, 8)) * (rng.random((6, 7, 8)) > 0.7)).astype(np.uint8), 1.0, 2),
        ("wide_r", (rng.random((16, 6, 120)) > 0.98).astype(np.uint8), 10.0, 2),
    ]
    out = {}
    for name, mask, sigma, reps in cases:
        k = gauss_kernel(sigma).astype(np.float32)
        u, q = scipy_diffusion(mask, k, reps)
        out[f"{name}__mask"] = mask
        out[f"{name}__kernel"] = k
        out[f"{name}__reps"] = np.int64(reps)
        out[f"{name}__u"] = u
        out[f"{name}__q"] = q
    np.savez_compressed(os.path.join(HERE, "small.npz"), names=np.array([c[0] for c in cases]), **out)

    # 4. asymmetric taps: correlation, not convolution
    out = {}
    names = []
    for name, shape, K, reps in [("asym5", (11, 13, 17), 5, 3), ("asym9", (19, 6, 40), 9, 2)]:
        mask = (rng.random(shape) > 0.93).astype(np.uint8)
        k = rng.random(K)
        k = (k / k.sum() * 0.97).astype(np.float32)
        u, q = scipy_diffusion(mask, k, reps, op=ndi.correlate1d)
        names.append(name)
        out[f"{name}__mask"] = mask
        out[f"{name}__kernel"] = k
        out[f"{name}__reps"] = np.int64(reps)
        out[f"{name}__u"] = u
        out[f"{name}__q"] = q
    np.savez_compressed(os.path.join(HERE, "correlate.npz"), names=np.array(names), **out)

    # 5. taps: impulse responses of gaussian_filter1d
    sigmas = [0.3, 0.8, 1.0, 1.7, 2.0, 3.0, 4.7, 10.0, 15.0, 7.3, 2.9]
    taps = {}
    for s in sigmas:
        R = int(4.0 * s + 0.5)
        assert R == round(4.0 * s), s
        d = np.zeros(2 * R + 1 + 40)
        d[R + 20] = 1.0
        resp = ndi.gaussian_filter1d(d, s, mode="constant", truncate=4.0)
        taps[f"s{s}"] = resp[20:20 + 2 * R + 1]
    np.savez_compressed(os.path.join(HERE, "taps.npz"), sigmas=np.array(sigmas), **taps)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
