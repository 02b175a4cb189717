"""Pin the oracle (CPU restatement of src/lib/cpp/cpu/diffusion.cc).

It is checked against (1) scipy known answers written by
tests/golden/make_golden.py (the reference test's own oracle,
src/test/test_diffusion.py:49-59), (2) the reference's own CPU module compiled
from its sources into oracle/_ref (bit-exact uint16), (3) an independent FP64
brute force on tiny volumes, and (4) closed forms and invariants.  Planted
mistakes in the oracle must fail these checks.
"""
import hashlib
import os
import tempfile

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from maxibone_amd.helpers import gauss_kernel, generate_cylinder_mask
from maxibone_amd.synth import WORKLOADS, kat_mask

F32_TOL_SCIPY = 1e-6  # scipy sums each line in double: differs by FP32 rounding only


# ---------------------------------------------------------------- taps
def test_gauss_kernel_matches_scipy_impulse_response():
    d = np.load(os.path.join(GOLDEN, "taps.npz"))
    for s in d["sigmas"]:
        want = d[f"s{s}"]
        for impl in (oracle.gauss_kernel, gauss_kernel):
            got = impl(float(s))
            assert got.shape == want.shape
            np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-17)


def test_gauss_kernel_is_bit_exact_with_the_oracle():
    # hlp:402-404 squares integer offsets and scales them by the FP64 factor
    # -0.5/sigma^2; the drop-in helper follows that order, so its FP64 taps
    # equal the oracle's bit for bit (not only to FP64 rounding)
    sigmas = np.concatenate([np.linspace(0.3, 20.0, 1970), [1.0, 2.2, 3.0, 10.0, 15.0]])
    for s in sigmas:
        a, b = gauss_kernel(float(s)), oracle.gauss_kernel(float(s))
        assert a.dtype == b.dtype == np.float64
        assert a.tobytes() == b.tobytes(), s


def test_gauss_kernel_bit_check_sees_operation_order():
    # the check above discriminates: scaling float offsets twice, the order a
    # restatement could plausibly take, changes FP64 bits for most sigmas
    def float_order(sigma):
        r = round(4.0 * sigma)
        x = np.arange(-r, r + 1, dtype=np.float64)
        phi = np.exp((-0.5 / (sigma * sigma)) * x * x)
        return phi / phi.sum()

    sigmas = np.linspace(0.3, 20.0, 200)
    differ = sum(float_order(float(s)).tobytes() != oracle.gauss_kernel(float(s)).tobytes()
                 for s in sigmas)
    assert differ > 50


@pytest.mark.parametrize("sigma,taps", [(3.0, 25), (10.0, 81), (15.0, 121), (1.125, 9), (0.625, 5),
                                        (1.375, 13), (0.375, 5), (0.875, 9)])
def test_gauss_kernel_radius_is_round_half_even(sigma, taps):
    # hlp:392 uses Python round(): 4.5 -> 4, 2.5 -> 2, 5.5 -> 6, 1.5 -> 2
    assert oracle.gauss_kernel(sigma).shape == (taps,)
    assert gauss_kernel(sigma).shape == (taps,)


def test_gauss_kernel_closed_form():
    k = gauss_kernel(2.5)
    R = 10
    x = np.arange(-R, R + 1)
    phi = np.exp(-x * x / (2 * 2.5 ** 2))
    np.testing.assert_allclose(k, phi / phi.sum(), rtol=1e-15)
    assert abs(k.sum() - 1.0) < 1e-15 and np.all(k == k[::-1])


def test_cylinder_mask_closed_form():
    m = generate_cylinder_mask(5)
    xs = np.linspace(-1, 1, 5)
    assert np.array_equal(m, (xs[None, :] ** 2 + xs[:, None] ** 2) < 1)
    assert generate_cylinder_mask(4, 6).shape == (4, 6)


# ---------------------------------------------------------------- fixtures
def _check_against_scipy(mask, k, reps, u_want, q_want):
    u, q = oracle.diffusion_f32(mask, k, reps)
    assert np.abs(u - u_want).max() <= F32_TOL_SCIPY
    assert np.abs(q.astype(np.int32) - q_want).max() <= 1  # tst:92-98


def test_oracle_small_fixtures(golden_small):
    for name, m, k, reps, u, q in golden_small:
        _check_against_scipy(m, k, reps, u, q)


def test_oracle_correlate_fixtures(golden_correlate):
    for name, m, k, reps, u, q in golden_correlate:
        _check_against_scipy(m, k, reps, u, q)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("fixture", ["kat_128", "c1_128"])
def test_oracle_sampled_fixtures(fixture):
    d = np.load(os.path.join(GOLDEN, f"{fixture}.npz"))
    m = kat_mask() if fixture == "kat_128" else WORKLOADS["c1"].mask()
    assert _sha(np.packbits(m.astype(bool))) == str(d["mask_sha"])
    u, q = oracle.diffusion_f32(m, d["kernel"], int(d["reps"]))
    n = m.shape[0]
    assert np.abs(q[n // 2].astype(np.int32) - d["slice_q"]).max() <= 1
    assert np.abs(q[::5, ::5, ::5].astype(np.int32) - d["sub_q"]).max() <= 1
    assert np.abs(u[::5, ::5, ::5] - d["sub_u"]).max() <= F32_TOL_SCIPY


# ---------------------------------------------------------------- reference build
@pytest.fixture(scope="module")
def ref():
    mod = oracle.ref_module()
    if mod is None:
        pytest.skip("reference CPU module not built (oracle/build.py)")
    return mod


@pytest.mark.parametrize("seed", range(6))
def test_oracle_bit_equal_to_reference_build(ref, seed):
    rng = np.random.default_rng(100 + seed)
    shape = tuple(int(s) for s in rng.integers(1, 40, 3))
    m = (rng.random(shape) > rng.uniform(0.8, 0.999)).astype(np.uint8)
    if seed % 2:
        k = rng.random(2 * int(rng.integers(0, 9)) + 1).astype(np.float32)
        k /= k.sum()
    else:
        k = gauss_kernel(float(rng.uniform(0.5, 6))).astype(np.float32)
    reps = int(rng.integers(0, 6))
    _, q = oracle.diffusion_f32(m, k, reps)
    q_ref = np.empty(shape, np.uint16)
    ref.diffusion(m, k, q_ref, reps)
    assert np.array_equal(q, q_ref)


@pytest.mark.parametrize("reps", [1, 3])
def test_oracle_chunked_quirk_equals_reference_on_disk(ref, reps):
    """The quirk-mode chunked oracle equals the reference's on-disk overload.

    The reference leaves its last fwrite to the state file in the stdio buffer
    (at most 4096 bytes) when convert_float_to_uint8 reopens that file by name
    (cpu:449-459), so the final <= 1024 floats of its output are stale; they
    are excluded here.
    """
    rng = np.random.default_rng(5)
    shape = (23, 9, 14)
    m = (rng.random(shape) > 0.9).astype(np.uint8)
    k = gauss_kernel(1.3).astype(np.float32)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.uint8"), os.path.join(td, "out.uint8")
        m.tofile(fin)
        ref.diffusion(fin, k, fout, shape, (7, 9, 14), reps)
        got = np.fromfile(fout, np.uint8, m.size).reshape(shape)
    u = oracle.diffusion_chunked_f32(m, k, reps, 7, quirk_mask=True)
    keep = m.size - 1024
    assert np.array_equal(oracle.quantise_u8(u).ravel()[:keep], got.ravel()[:keep])


# ---------------------------------------------------------------- brute force FP64
def _conv_matrix(n, k):
    R = len(k) // 2
    M = np.zeros((n, n))
    for i in range(n):
        for r in range(-R, R + 1):
            if 0 <= i + r < n:
                M[i, i + r] = k[R + r]
    return M


def _brute_f64(mask, k, reps):
    a = mask > 0
    u = a.astype(np.float64)
    Mz, My, Mx = (_conv_matrix(n, k) for n in mask.shape)
    for _ in range(reps):
        u = np.einsum("ij,jkl->ikl", Mz, u)
        u = np.einsum("ij,kjl->kil", My, u)
        u = np.einsum("ij,klj->kli", Mx, u)
        u[a] = 1.0
    return u


@pytest.mark.parametrize("shape,K,reps", [((5, 6, 7), 3, 3), ((9, 4, 11), 7, 2), ((3, 3, 3), 9, 4),
                                          ((12, 1, 5), 5, 3)])
def test_oracle_fp64_matches_brute_force(shape, K, reps):
    rng = np.random.default_rng(sum(shape) + K)
    m = (rng.random(shape) > 0.8).astype(np.uint8)
    k = rng.random(K)
    k /= k.sum()
    want = _brute_f64(m, k, reps)
    np.testing.assert_allclose(oracle.diffusion_f64(m, k, reps), want, rtol=0, atol=1e-13)
    u32, _ = oracle.diffusion_f32(m, k.astype(np.float32), reps)
    assert np.abs(u32 - want).max() < 1e-6


# ---------------------------------------------------------------- closed forms / invariants
def test_impulse_response_is_separable_product():
    m = np.zeros((25, 25, 25), np.uint8)
    m[12, 12, 12] = 1
    k = gauss_kernel(3.0)
    u = oracle.diffusion_f64(m, k, 1)
    want = np.einsum("i,j,l->ijl", k, k, k)
    want[12, 12, 12] = 1.0
    np.testing.assert_allclose(u, want, rtol=1e-13, atol=1e-18)


def test_pass_linearity_and_zero_boundary():
    rng = np.random.default_rng(3)
    x, y = rng.random((2, 8, 9, 10))
    k = gauss_kernel(1.5)
    for dim in range(3):
        lhs = oracle.core_f64(2.5 * x + y, k, dim)
        rhs = 2.5 * oracle.core_f64(x, k, dim) + oracle.core_f64(y, k, dim)
        np.testing.assert_allclose(lhs, rhs, rtol=1e-13)
    # the first voxel of a constant line sums only the in-range taps
    one = np.ones((1, 1, 20))
    R = len(k) // 2
    assert abs(oracle.core_f64(one, k, 2)[0, 0, 0] - k[R:].sum()) < 1e-15


def test_invariants_bounds_mask_monotone():
    m = WORKLOADS["c1"].mask()[:48, :40, :56].copy()
    k = gauss_kernel(2.0).astype(np.float32)
    u_prev = (m > 0).astype(np.float32)
    for reps in (1, 2, 3):
        u, q = oracle.diffusion_f32(m, k, reps)
        assert u.min() >= 0 and u.max() <= 1
        assert np.all(u[m > 0] == 1.0)
        assert np.all(u >= u_prev - 1e-7)  # G >= 0 and illumination: non-decreasing
        u_prev = u
    u0, q0 = oracle.diffusion_f32(m, k, 0)
    assert np.array_equal(q0, (m > 0).astype(np.uint16) * 65535)


def test_quantise_closed_form():
    u = np.array([0, 1, 0.5, 1 / 65535, np.nextafter(np.float32(1 / 65535), np.float32(0)), 0.99999],
                 np.float32)
    want = np.floor(u * np.float32(65535)).astype(np.uint16)
    assert np.array_equal(oracle.quantise_u16(u), want)
    assert oracle.quantise_u16(np.float32([1.0]))[0] == 65535


def test_dyadic_taps_keep_two_repetitions_exact():
    """The premise of the GPU test of the 16-bit forms' final repetition: with
    taps 1/4, 1/2, 1/4 and a 0/1 start, the first repetition's values are
    multiples of 1/64 with at most 7 significant bits (exact in BF16), and two
    repetitions stay exact in FP32 (the FP32 oracle equals the FP64 one bit
    for bit), so every arithmetic form must return the FP32 oracle's bits."""
    rng = np.random.default_rng(2)
    k = np.array([0.25, 0.5, 0.25], np.float32)
    for shape in [(20, 24, 64), (9, 40, 136), (1, 1, 64), (3, 1, 8)]:
        m = (rng.random(shape) > 0.9).astype(np.uint8)
        u1, _ = oracle.diffusion_f32(m, k, 1)
        scaled = u1.astype(np.float64) * 64
        assert np.array_equal(scaled, np.round(scaled))
        bf16 = (u1.view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32)  # truncation to BF16
        assert np.array_equal(bf16, u1)
        u2, _ = oracle.diffusion_f32(m, k, 2)
        assert np.array_equal(u2.astype(np.float64), oracle.diffusion_f64(m, k, 2))
    # a Gaussian's taps are not dyadic: there the two oracles differ
    k = gauss_kernel(1.0).astype(np.float32)
    m = (rng.random((20, 24, 64)) > 0.9).astype(np.uint8)
    assert not np.array_equal(oracle.diffusion_f32(m, k, 2)[0].astype(np.float64), oracle.diffusion_f64(m, k, 2))


@pytest.mark.parametrize("gz", [1, 4, 7, 30])
def test_chunked_true_mask_equals_in_memory(gz):
    rng = np.random.default_rng(gz)
    m = (rng.random((29, 11, 13)) > 0.93).astype(np.uint8)
    k = gauss_kernel(1.7).astype(np.float32)
    u, _ = oracle.diffusion_f32(m, k, 4)
    assert np.array_equal(oracle.diffusion_chunked_f32(m, k, 4, gz), u)


# ---------------------------------------------------------------- planted mistakes
def test_planted_mistakes_fail_the_fixtures(golden_correlate, golden_small):
    """Each plausible slip, planted in the oracle's inputs, must break a check."""
    name, m, k, reps, u_want, _ = golden_correlate[0]
    flipped, _ = oracle.diffusion_f32(m, k[::-1].copy(), reps)  # convolution instead of correlation
    assert np.abs(flipped - u_want).max() > 1e-3
    shifted = np.roll(k, 1)  # off-by-one tap index
    assert np.abs(oracle.diffusion_f32(m, shifted, reps)[0] - u_want).max() > 1e-3
    name, m, k, reps, u_want, _ = [c for c in golden_small if c[0] == "mid"][0]
    no_illum = oracle.diffusion_f32(m, k, reps)[0].copy()
    fewer = oracle.diffusion_f32(m, k, reps - 1)[0]  # one repetition dropped
    assert np.abs(fewer - u_want).max() > 1e-3
    two_pass = oracle.core_f32(oracle.core_f32((m > 0).astype(np.float32), k, 0), k, 1)  # x pass dropped
    one_rep = oracle.diffusion_f32(m, k, 1)[0]
    assert np.abs(np.where(m > 0, 1, two_pass) - one_rep).max() > 1e-3
    assert np.abs(no_illum - u_want).max() <= F32_TOL_SCIPY
