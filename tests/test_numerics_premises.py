"""Numerical premises of the device arithmetic, checked on the CPU in numpy.

The kernels rely on two facts about IEEE rounding that no GPU test could
isolate: (1) for a power-of-two scale, the low FP16 part of a split formed by
one rounding of the exact a * scale - hi (an FMA) equals the two-step form
fp16(fp32(a * scale) - hi); (2) a BF16 value widened to FP16 is exact from
2^-14 (FP16's smallest normal number) up, and within 2^-25 below.
"""
import numpy as np


def _split_two_step(a, scale):
    v = (a * scale).astype(np.float32)
    hi = v.astype(np.float16)
    lo = (v - hi.astype(np.float32)).astype(np.float16)
    return hi, lo


def _split_fma(a, scale):
    v = (a * scale).astype(np.float32)
    hi = v.astype(np.float16)
    exact = a.astype(np.float64) * float(scale) - hi.astype(np.float64)  # exact in FP64
    return hi, exact.astype(np.float16)  # one rounding, as the FMA's


def test_fma_low_part_equals_the_two_step_split():
    rng = np.random.default_rng(0)
    # accumulators of taps scaled into [2^14, 2^15): sums up to a few times 2^15
    a = np.concatenate([rng.uniform(0, 2 ** 16, 200000), rng.uniform(0, 1e-3, 50000),
                        np.float32([0.0, 1.0, 2 ** 15, 65535.0])]).astype(np.float32)
    for s in (14, 15, 16, 20):
        scale = np.float32(2.0 ** -s)
        h1, l1 = _split_two_step(a, scale)
        h2, l2 = _split_fma(a, scale)
        assert np.array_equal(h1.view(np.uint16), h2.view(np.uint16))
        assert np.array_equal(l1.view(np.uint16), l2.view(np.uint16))


def test_the_fma_check_sees_a_double_rounding():
    """The same comparison with a scale that is not a power of two (the
    product then rounds before the subtraction) finds differences."""
    rng = np.random.default_rng(1)
    a = rng.uniform(0, 2 ** 16, 200000).astype(np.float32)
    scale = np.float32(1.0 / 3.0)
    _, l1 = _split_two_step(a, scale)
    _, l2 = _split_fma(a, scale)
    assert not np.array_equal(l1.view(np.uint16), l2.view(np.uint16))


def test_bf16_widens_to_fp16():
    # every BF16 bit pattern of a finite value in [0, 65504]
    bits = np.arange(0, 0x7F80, dtype=np.uint32)
    v = (bits << 16).view(np.float32)
    v = v[v <= 65504.0]
    h = v.astype(np.float16).astype(np.float32)
    normal = v >= 2.0 ** -14
    assert np.array_equal(h[normal], v[normal])
    assert np.abs(h[~normal] - v[~normal]).max() <= 2.0 ** -25
