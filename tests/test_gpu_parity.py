"""Parity of the HIP path with the oracle and the golden fixtures (needs a GPU).

Bars (FP32 state):
* precision "fp32_exact": bit-identical to the oracle (the reference CPU
  arithmetic: r-ascending sums, product rounded then added).
* precision "fp32" (MFMA), "fp32_fma" and "fp32_split" (FP16 hi + lo state,
  three FP16 MFMAs per chunk): max |u - u_oracle| <= 1e-5 and uint16 within
  1 LSB (src/test/test_diffusion.py:92-98).
* drop-in `diffusion` (uint16): within 1 LSB of the oracle and of the scipy
  known answers.
"""
import os
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, load_cases
from maxibone_amd.helpers import gauss_kernel
from maxibone_amd.synth import WORKLOADS, kat_mask

pytestmark = pytest.mark.gpu
F32_TOL = 1e-5


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _u16_close(a, b):
    return np.abs(a.astype(np.int32) - b.astype(np.int32)).max(initial=0) <= 1


def _check_f32(gpu, m, k, reps, prec):
    u_ref, q_ref = oracle.diffusion_f32(m, k, reps)
    q, u = gpu.diffusion(m, k, reps, precision=prec, u16=True, f32=True)
    if prec == "fp32_exact":
        assert np.array_equal(u.view(np.uint32), u_ref.view(np.uint32)), "not bit-identical"
        assert np.array_equal(q, q_ref)
    else:
        err = np.abs(u - u_ref).max(initial=0)
        assert err <= F32_TOL, f"max-abs {err}"
        assert _u16_close(q, q_ref)
    return u, q


@pytest.mark.parametrize("prec", ["fp32_exact", "fp32", "fp32_fma", "fp32_split"])
def test_small_fixtures(gpu, golden_small, golden_correlate, prec):
    for name, m, k, reps, u_sc, q_sc in golden_small + golden_correlate:
        _say("case", name, m.shape, len(k), reps, prec)
        u, q = _check_f32(gpu, m, k, reps, prec)
        assert _u16_close(q, q_sc)


def test_drop_in_module_matches_fixtures(gpu, golden_small):
    from maxibone_amd.diffusion import diffusion
    for name, m, k, reps, u_sc, q_sc in golden_small:
        _say("drop-in", name)
        out = np.zeros(m.shape, np.uint16)
        diffusion(m, k, out, reps)
        assert _u16_close(out, q_sc)
        _, q_ref = oracle.diffusion_f32(m, k, reps)
        assert _u16_close(out, q_ref)
    # bool masks and float64 kernels are cast like the reference's forcecast
    name, m, k, reps, u_sc, q_sc = golden_small[6]
    out = np.zeros(m.shape, np.uint16)
    diffusion(m.astype(bool), k.astype(np.float64), out, repititions=reps, verbose=0)
    assert _u16_close(out, q_sc)


RADII = [1, 2, 3, 5, 6, 9, 12, 13, 17, 21, 27, 31, 36, 40, 44, 52, 60, 63, 64, 65, 90]


@pytest.mark.parametrize("R", RADII)
def test_radius_sweep(gpu, R):
    rng = np.random.default_rng(R)
    shape = (int(rng.integers(20, 50)), int(rng.integers(20, 70)), int(rng.integers(60, 200)))
    m = (rng.random(shape) > 0.995).astype(np.uint8)
    k = gauss_kernel(R / 4.0).astype(np.float32)
    assert len(k) == 2 * R + 1
    for prec in ("fp32_exact", "fp32", "fp32_fma", "fp32_split"):
        _say("R", R, shape, prec)
        _check_f32(gpu, m, k, 2, prec)


SHAPES = [(1, 1, 1), (1, 1, 70), (70, 1, 1), (1, 70, 1), (2, 3, 4), (65, 2, 63), (3, 129, 17),
          (17, 64, 64), (33, 65, 130), (130, 5, 9)]


@pytest.mark.parametrize("shape", SHAPES)
def test_edge_shapes(gpu, shape):
    rng = np.random.default_rng(sum(shape))
    m = (rng.random(shape) > 0.9).astype(np.uint8)
    for sigma in (1.0, 3.0, 6.0):
        k = gauss_kernel(sigma).astype(np.float32)
        for prec in ("fp32_exact", "fp32", "fp32_split"):
            _say("shape", shape, sigma, prec)
            _check_f32(gpu, m, k, 3, prec)


def test_asymmetric_taps(gpu):
    rng = np.random.default_rng(11)
    m = (rng.random((30, 40, 90)) > 0.97).astype(np.uint8)
    for K in (3, 11, 25, 81):
        k = rng.random(K).astype(np.float32)
        k /= k.sum() * 1.01
        for prec in ("fp32_exact", "fp32", "fp32_fma", "fp32_split"):
            _say("asym", K, prec)
            _check_f32(gpu, m, k, 3, prec)


@pytest.mark.parametrize("fixture", ["kat_128", "c1_128"])
def test_sampled_fixtures(gpu, fixture):
    d = np.load(os.path.join(GOLDEN, f"{fixture}.npz"))
    m = kat_mask().astype(np.uint8) if fixture == "kat_128" else WORKLOADS["c1"].mask()
    reps = int(d["reps"])
    _say("fixture", fixture, reps)
    u_ref, q_ref = oracle.diffusion_f32(m, d["kernel"], reps)
    from maxibone_amd.diffusion import diffusion
    out = np.zeros(m.shape, np.uint16)
    diffusion(m, d["kernel"], out, reps)
    n = m.shape[0]
    assert _u16_close(out[n // 2], d["slice_q"])
    assert _u16_close(out[::5, ::5, ::5], d["sub_q"])
    assert _u16_close(out, q_ref)
    for prec in ("fp32", "fp32_exact"):
        q, u = gpu.diffusion(m, d["kernel"], reps, precision=prec, f32=True)
        if prec == "fp32_exact":
            assert np.array_equal(u, u_ref)
        else:
            assert np.abs(u - u_ref).max() <= F32_TOL


@pytest.mark.parametrize("chunk_z,temporal", [(7, 1), (16, 1), (16, 3), (40, 2), (100, 1), (5, 4)])
def test_chunked_equals_in_memory(gpu, chunk_z, temporal):
    rng = np.random.default_rng(chunk_z * 10 + temporal)
    m = (rng.random((61, 33, 70)) > 0.97).astype(np.uint8)
    k = gauss_kernel(2.2).astype(np.float32)
    for prec in ("fp32_exact", "fp32", "fp32_split"):
        _say("chunked", chunk_z, temporal, prec)
        q0, u0 = gpu.diffusion(m, k, 5, precision=prec, f32=True)
        q1, u1 = gpu.diffusion_chunked(m, k, 5, chunk_z, temporal, precision=prec, f32=True)
        assert np.array_equal(u0, u1) and np.array_equal(q0, q1)
    u_ref = oracle.diffusion_chunked_f32(m, k, 5, chunk_z)
    assert np.array_equal(gpu.diffusion_chunked(m, k, 5, chunk_z, temporal, "fp32_exact", f32=True)[1], u_ref)


@pytest.mark.parametrize("reps,temporal", [(0, 1), (3, 3), (3, 5), (4, 1), (7, 3)])
def test_chunked_single_block_and_outputs(gpu, reps, temporal):
    """One block of repetitions (reps <= temporal: no host state), zero
    repetitions, and each output alone, against the in-memory path."""
    rng = np.random.default_rng(40 + reps * 7 + temporal)
    m = (rng.random((29, 21, 40)) > 0.96).astype(np.uint8)
    k = gauss_kernel(1.6).astype(np.float32)
    _say("chunked outputs", reps, temporal)
    q0, u0 = gpu.diffusion(m, k, reps, f32=True)
    q1, u1 = gpu.diffusion_chunked(m, k, reps, 8, temporal, f32=True)
    assert np.array_equal(u0, u1) and np.array_equal(q0, q1)
    q2, _ = gpu.diffusion_chunked(m, k, reps, 8, temporal, u16=True, f32=False)
    assert np.array_equal(q2, q0)
    _, u3 = gpu.diffusion_chunked(m, k, reps, 8, temporal, u16=False, f32=True)
    assert np.array_equal(u3, u0)


def test_drop_in_large_volume_staged_copies(gpu):
    """Host buffers of 128 MB and more travel through the pinned staging
    chunks (mask in, uint16 and float32 out, a ragged last chunk): the drop-in
    call equals the device-resident entry bit for bit."""
    from maxibone_amd import hip
    rng = np.random.default_rng(77)
    shape = (520, 512, 512)
    m = np.zeros(shape, np.uint8)
    m[rng.random(shape) > 0.9995] = 1
    m[:, 200:260, 300:330] = 1
    k = gauss_kernel(1.0).astype(np.float32)
    _say("staged copies", shape)
    q, u = gpu.diffusion(m, k, 1, f32=True)
    L = gpu.lib()
    sb = L.mxb_state_bytes(gpu.PRECISION["fp32"], *shape)
    dm = hip.DeviceBuffer.from_array(m)
    s0, s1 = hip.DeviceBuffer(sb), hip.DeviceBuffer(sb)
    dq, df = hip.DeviceBuffer(2 * m.size), hip.DeviceBuffer(4 * m.size)
    gpu.check(L.mxb_device_diffusion(dm.ptr, *shape, k.ctypes.data, len(k), 1, gpu.PRECISION["fp32"],
                                     s0.ptr, s1.ptr, dq.ptr, df.ptr, None))
    hip.synchronize()
    assert np.array_equal(q, dq.to_array(np.uint16, shape))
    assert np.array_equal(u, df.to_array(np.float32, shape))
    for b in (dm, s0, s1, dq, df):
        b.free()


def test_on_disk_overload(gpu, tmp_path):
    from maxibone_amd.diffusion import diffusion
    rng = np.random.default_rng(21)
    shape = (45, 20, 36)
    m = (rng.random(shape) > 0.95).astype(np.uint8)
    k = gauss_kernel(1.8).astype(np.float32)
    fin, fout = str(tmp_path / "mask.uint8"), str(tmp_path / "out.uint16")
    m.tofile(fin)
    _say("on-disk")
    diffusion(fin, k, fout, shape, (10, 20, 36), 4)
    got = np.fromfile(fout, np.uint16).reshape(shape)
    want = np.zeros(shape, np.uint16)
    diffusion(m, k, want, 4)
    assert np.array_equal(got, want)
    _, q_ref = oracle.diffusion_f32(m, k, 4)
    assert _u16_close(got, q_ref)


def _batch_inputs(seed, n, lo, hi, s_lo, s_hi):
    rng = np.random.default_rng(seed)
    masks, kernels = [], []
    for i in range(n):
        shape = tuple(int(s) for s in rng.integers(lo, hi, 3))
        masks.append((rng.random(shape) > 0.97).astype(np.uint8))
        kernels.append(gauss_kernel(float(rng.uniform(s_lo, s_hi))).astype(np.float32))
    return masks, kernels


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_batched_host_entry(gpu, three_pass, prec):
    masks, kernels = _batch_inputs(31, 6, 5, 70, 0.7, 10.0)
    _say("batched host", prec)
    qs, fs = gpu.diffusion_batched(masks, kernels, 3, precision=prec, f32=True)
    for m, k, q, f in zip(masks, kernels, qs, fs):
        if prec == "fp32":
            u_ref, q_ref = oracle.diffusion_f32(m, k, 3)
            # the FP32 MFMA chain equals the FMA chain: bit-equal to single-volume runs
            single = gpu.diffusion(m, k, 3, precision="fp32", f32=True)[1]
            assert np.array_equal(f, single)
            assert np.abs(f - u_ref).max() <= F32_TOL and _u16_close(q, q_ref)
        else:
            # the persistent 16-bit passes equal the single-volume 16-bit
            # path in its three-pass form
            q1, f1 = gpu.diffusion(m, k, 3, precision=prec, u16=True, f32=True)
            assert np.array_equal(f, f1) and np.array_equal(q, q1)
            u64 = oracle.diffusion_f64(m, k.astype(np.float64), 3)
            assert np.abs(f - u64).max() <= LOWP_TOL


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_batched_plan_persistent_graph(gpu, prec):
    """The captured persistent graph equals per-volume launches, on every replay."""
    from maxibone_amd import hip
    masks, kernels = _batch_inputs(32, 9, 20, 150, 1.0, 10.0)
    dm = [hip.DeviceBuffer.from_array(m) for m in masks]
    of = [hip.DeviceBuffer(4 * m.size) for m in masks]
    oq = [hip.DeviceBuffer(2 * m.size) for m in masks]
    plan = gpu.BatchPlan([b.ptr for b in dm], [m.shape for m in masks], kernels, 4,
                         outs_u16=[b.ptr for b in oq], outs_f32=[b.ptr for b in of], precision=prec)
    info = plan.info()
    _say("batch plan", info)
    assert info["persistent"] and info["grid"] > 0
    plan.run(1)
    plan.sync()
    want = [b.to_array(np.float32, m.shape) for b, m in zip(of, masks)]
    wantq = [b.to_array(np.uint16, m.shape) for b, m in zip(oq, masks)]
    for rep in range(3):
        for b in of + oq:
            b.fill_bytes(0xFF)  # NaN / 65535 until the graph writes
        hip.synchronize()
        plan.run(0)
        plan.sync()
        for b, w, bq, wq, m in zip(of, want, oq, wantq, masks):
            assert np.array_equal(b.to_array(np.float32, m.shape), w)
            assert np.array_equal(bq.to_array(np.uint16, m.shape), wq)
    for m, k, f in zip(masks, kernels, want):
        if prec == "fp32":
            u_ref, _ = oracle.diffusion_f32(m, k, 4)
            assert np.abs(f - u_ref).max() <= F32_TOL
        else:
            u64 = oracle.diffusion_f64(m, k.astype(np.float64), 4)
            assert np.abs(f - u64).max() <= LOWP_TOL
    plan.close()


def test_zero_repetitions_and_determinism(gpu):
    m = WORKLOADS["c1"].mask()
    k = gauss_kernel(3.0).astype(np.float32)
    q0, u0 = gpu.diffusion(m, k, 0, f32=True)
    assert np.array_equal(q0, (m > 0).astype(np.uint16) * 65535)
    a = gpu.diffusion(m, k, 4, f32=True)[1]
    b = gpu.diffusion(m, k, 4, f32=True)[1]
    assert np.array_equal(a, b)


# ---------------------------------------------------------------- precision form
LOWP_TOL = 2e-3  # max-abs vs the FP64 oracle (BF16/FP16 storage, FP32-accuracy last repetition)


def _lowp_errors(gpu, m, k32, reps, prec):
    u64 = oracle.diffusion_f64(m, k32.astype(np.float64), reps)
    q, u = gpu.diffusion(m, k32, reps, precision=prec, f32=True)
    d = u.astype(np.float64) - u64
    return np.abs(d).max(), np.sqrt(np.mean(d * d)), q, u


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
@pytest.mark.parametrize("sigma,reps", [(1.0, 6), (3.0, 20), (6.0, 10), (10.0, 8)])
def test_precision_form_error_vs_fp64(gpu, prec, sigma, reps):
    m = WORKLOADS["c1"].mask()[:80, :96, :112].copy()
    k = gauss_kernel(sigma).astype(np.float32)
    mx, rms, q, u = _lowp_errors(gpu, m, k, reps, prec)
    _say("lowp", prec, sigma, reps, "max", mx, "rms", rms)
    assert mx <= LOWP_TOL
    assert np.all(u[m > 0] == 1.0)
    assert u.min() >= 0.0 and u.max() <= 1.0 + 1e-6


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_precision_form_edges(gpu, prec):
    rng = np.random.default_rng(41)
    for shape in [(1, 1, 1), (5, 3, 70), (37, 2, 19), (33, 65, 130)]:
        m = (rng.random(shape) > 0.9).astype(np.uint8)
        for sigma in (0.7, 2.5, 12.0, 16.0):
            k = gauss_kernel(sigma).astype(np.float32)
            _say("lowp edge", prec, shape, sigma)
            mx, rms, _, _ = _lowp_errors(gpu, m, k, 4, prec)
            assert mx <= LOWP_TOL


def test_unaligned_rows_direct_load(gpu):
    """Rows of 104 voxels (416 / 208 bytes, not whole 128-byte lines): the
    direct-load passes store them with the plain policy, in the XCD-contiguous
    tile order, for FP32 and both 16-bit forms."""
    rng = np.random.default_rng(104)
    m = (rng.random((40, 40, 104)) > 0.95).astype(np.uint8)
    k = gauss_kernel(4.0).astype(np.float32)
    _say("unaligned rows", m.shape, "fp32")
    _check_f32(gpu, m, k, 4, "fp32")
    for prec in ("bf16", "fp16"):
        _say("unaligned rows", m.shape, prec)
        mx, _, _, _ = _lowp_errors(gpu, m, k, 6, prec)
        assert mx <= LOWP_TOL


# planes that are whole 64-lane groups with rows that are not: the z-pass runs
# each plane as one run of ny * nx lanes
FLAT_Z_SHAPES = [(24, 48, 100), (9, 64, 125), (13, 16, 1000), (6, 192, 40)]


@pytest.mark.parametrize("prec", ["fp32_exact", "fp32", "fp32_split", "bf16"])
def test_flat_plane_z_pass(gpu, prec):
    """Volumes whose z-pass takes the flat plane layout: the exact form stays
    bit-identical to the oracle, the others within their bars."""
    rng = np.random.default_rng(64)
    for shape in FLAT_Z_SHAPES:
        m = (rng.random(shape) > 0.94).astype(np.uint8)
        for sigma in (1.0, 5.0):
            k = gauss_kernel(sigma).astype(np.float32)
            _say("flat z", prec, shape, sigma)
            if prec == "bf16":
                mx, _, _, _ = _lowp_errors(gpu, m, k, 5, prec)
                assert mx <= LOWP_TOL
            else:
                _check_f32(gpu, m, k, 5, prec)


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_precision_form_final_repetition_on_exact_values(gpu, prec):
    """Taps 1/4, 1/2, 1/4 keep two repetitions from a 0/1 state exact: the
    first repetition's values (multiples of 1/64) are exact 16-bit numbers,
    and the final one, which starts from that 16-bit state (in the split
    form's widening z-pass and fused pass where the fused pass serves the
    volume, else in the FP32 form), is exact as well.  So the result equals
    the FP32 oracle bit for bit, in-memory and chunked (where a chunk visit
    runs the final repetition alone).  Shapes: standard and flat-plane z
    layouts, single rows and planes, and rows of 100 voxels, which the fused
    pass does not serve."""
    rng = np.random.default_rng(2)
    k = np.array([0.25, 0.5, 0.25], np.float32)
    for shape in [(20, 24, 64), (9, 40, 136), (16, 8, 1000), (1, 1, 64), (3, 1, 8), (1, 40, 16), (6, 30, 100)]:
        m = (rng.random(shape) > 0.9).astype(np.uint8)
        u_ref, q_ref = oracle.diffusion_f32(m, k, 2)
        _say("exact final repetition", prec, shape)
        q, u = gpu.diffusion(m, k, 2, precision=prec, f32=True)
        assert np.array_equal(u.view(np.uint32), u_ref.view(np.uint32)) and np.array_equal(q, q_ref)
        qc, uc = gpu.diffusion_chunked(m, k, 2, 5, 1, precision=prec, f32=True)
        assert np.array_equal(uc.view(np.uint32), u_ref.view(np.uint32)) and np.array_equal(qc, q_ref)


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_precision_form_one_repetition_is_fp32(gpu, prec):
    m = WORKLOADS["c1"].mask()[:40, :50, :60].copy()
    k = gauss_kernel(2.0).astype(np.float32)
    a = gpu.diffusion(m, k, 1, precision=prec, f32=True)[1]
    b = gpu.diffusion(m, k, 1, precision="fp32", f32=True)[1]
    assert np.array_equal(a, b)
    q0 = gpu.diffusion(m, k, 0, precision=prec)[0]
    assert np.array_equal(q0, (m > 0).astype(np.uint16) * 65535)


def test_precision_form_rejects_large_radius(gpu):
    m = np.zeros((4, 4, 4), np.uint8)
    with pytest.raises(ValueError):
        gpu.diffusion(m, np.ones(2 * 70 + 1, np.float32) / 141, 3, precision="bf16")


def test_mfma_chain_equals_fma_chain(gpu):
    """The FP32 MFMA form accumulates taps in order (one chain per output), so
    it agrees bit for bit with the one-FMA-per-tap VALU form, whatever the
    tile alignment (in-memory vs chunked)."""
    rng = np.random.default_rng(77)
    m = (rng.random((48, 40, 72)) > 0.97).astype(np.uint8)
    for sigma in (1.5, 4.0, 10.0):
        k = gauss_kernel(sigma).astype(np.float32)
        a = gpu.diffusion(m, k, 3, precision="fp32", u16=False, f32=True)[1]
        b = gpu.diffusion(m, k, 3, precision="fp32_fma", u16=False, f32=True)[1]
        _say("mfma vs fma", sigma, np.abs(a - b).max())
        assert np.array_equal(a, b)


@pytest.fixture
def three_pass():
    """MXB_XY_FUSED=0 (read by the library at every call): the 16-bit and
    split forms run their three-pass structure, the one the batched
    persistent passes share; restores the default afterwards."""
    old = os.environ.get("MXB_XY_FUSED")
    os.environ["MXB_XY_FUSED"] = "0"
    yield
    if old is None:
        os.environ.pop("MXB_XY_FUSED", None)
    else:
        os.environ["MXB_XY_FUSED"] = old


@pytest.fixture
def pass_kernel():
    """Select the line-pass kernel form through MXB_PASS_KERNEL (read by the
    library at every launch); restores the default afterwards."""
    old = os.environ.get("MXB_PASS_KERNEL")

    def use(form):
        if form:
            os.environ["MXB_PASS_KERNEL"] = form
        else:
            os.environ.pop("MXB_PASS_KERNEL", None)

    yield use
    use(old)


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp32"])
def test_kernel_forms_agree(gpu, pass_kernel, three_pass, prec):
    """The register-ring, VGPR-staged LDS and direct-load LDS forms of a pass
    group taps into MFMA chunks identically, so they agree bit for bit.
    Shapes cover segment tails (axes not a multiple of 32), partial 64-line
    groups, and line strides that route the illuminating pass to another
    form; the direct-load form runs three times (its pipeline must not race)."""
    rng = np.random.default_rng(5)
    shapes = [(64, 64, 128), (37, 40, 208), (33, 24, 72), (20, 8, 48)]
    for shape in shapes:
        m = (rng.random(shape) > 0.93).astype(np.uint8)
        for sigma in (1.2, 5.0, 10.0):
            k = gauss_kernel(sigma).astype(np.float32)
            outs = {}
            for form in ("ring", "staged", ""):
                pass_kernel(form)
                outs[form] = gpu.diffusion(m, k, 4, precision=prec, u16=False, f32=True)[1]
            again = [gpu.diffusion(m, k, 4, precision=prec, u16=False, f32=True)[1] for _ in range(2)]
            pass_kernel(None)
            _say("forms", prec, shape, sigma,
                 {f or "direct": float(np.abs(o - outs["ring"]).max()) for f, o in outs.items()})
            for f, o in outs.items():
                assert np.array_equal(o, outs["ring"]), f"form {f or 'direct'} differs"
            for o in again:
                assert np.array_equal(o, outs[""])


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_kernel_forms_agree_random(gpu, pass_kernel, three_pass, prec):
    """Seeded random shapes and radii (1..64): the default (direct-load) form
    equals the register-ring form bit for bit, twice in a row."""
    rng = np.random.default_rng(2026 if prec == "fp32" else 2027)
    for case in range(10):
        shape = tuple(int(v) for v in rng.integers(1, 90, 3))
        if case % 3 == 0:  # aligned line counts exercise the direct-load forms
            shape = (shape[0], 8 * max(1, shape[1] // 8), 16 * max(1, shape[2] // 16))
        sigma = float(rng.uniform(0.3, 16.0))
        reps = int(rng.integers(1, 5))
        m = (rng.random(shape) > rng.uniform(0.8, 0.99)).astype(np.uint8)
        k = gauss_kernel(sigma).astype(np.float32)
        if k.size // 2 > 64:
            k = gauss_kernel(15.0).astype(np.float32)
        pass_kernel("ring")
        want = gpu.diffusion(m, k, reps, precision=prec, u16=True, f32=True)
        pass_kernel(None)
        got = [gpu.diffusion(m, k, reps, precision=prec, u16=True, f32=True) for _ in range(2)]
        _say("random forms", prec, shape, round(sigma, 2), reps)
        for q, u in got:
            assert np.array_equal(u, want[1]) and np.array_equal(q, want[0])


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_batched_random_shapes_equal_single_volume(gpu, three_pass, prec):
    """Batched plans over random shapes and radii equal the single-volume path
    (its three-pass form for the 16-bit precisions)."""
    rng = np.random.default_rng(31)
    masks, kernels = [], []
    for _ in range(7):
        shape = tuple(int(v) for v in rng.integers(3, 70, 3))
        masks.append((rng.random(shape) > 0.95).astype(np.uint8))
        kernels.append(gauss_kernel(float(rng.uniform(0.5, 12.0))).astype(np.float32))
    _say("batched random shapes", prec)
    qs, fs = gpu.diffusion_batched(masks, kernels, 3, precision=prec, u16=True, f32=True)
    for m, k, q, f in zip(masks, kernels, qs, fs):
        q1, f1 = gpu.diffusion(m, k, 3, precision=prec, u16=True, f32=True)
        assert np.array_equal(f, f1) and np.array_equal(q, q1)


# ---------------------------------------------------------------- streaming paths
@pytest.mark.parametrize("prec", ["bf16", "fp16"])
@pytest.mark.parametrize("chunk_z,temporal", [(7, 1), (16, 3), (40, 2), (5, 4)])
def test_chunked_16bit_equals_in_memory(gpu, prec, chunk_z, temporal):
    """16-bit chunked runs keep 16-bit host state between chunk visits and run
    the last repetition at FP32 accuracy from it (T = 1: a chunk visit runs
    it alone): bit-equal to the in-memory 16-bit form."""
    rng = np.random.default_rng(chunk_z * 31 + temporal)
    m = (rng.random((77, 33, 70)) > 0.97).astype(np.uint8)
    k = gauss_kernel(2.2).astype(np.float32)
    for reps in (1, 2, 6):
        _say("chunked 16-bit", prec, chunk_z, temporal, reps)
        q0, u0 = gpu.diffusion(m, k, reps, precision=prec, f32=True)
        q1, u1 = gpu.diffusion_chunked(m, k, reps, chunk_z, temporal, precision=prec, f32=True)
        assert np.array_equal(u0.view(np.uint32), u1.view(np.uint32)) and np.array_equal(q0, q1)


@pytest.fixture
def device_budget():
    """Set MXB_DEVICE_BUDGET (read by the library at every call); restores it."""
    old = os.environ.get("MXB_DEVICE_BUDGET")

    def use(nbytes):
        if nbytes is None:
            os.environ.pop("MXB_DEVICE_BUDGET", None)
        else:
            os.environ["MXB_DEVICE_BUDGET"] = str(int(nbytes))

    yield use
    use(old)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp32_exact"])
def test_drop_in_streams_when_the_volume_exceeds_the_budget(gpu, device_budget, prec):
    """A volume larger than the device budget takes the z-chunk streaming path
    (the reference switches to its out-of-core form, pyb:40-45): results are
    bit-equal to the in-HBM form, for the drop-in module and the C ABI."""
    from maxibone_amd.diffusion import diffusion
    rng = np.random.default_rng(404)
    m = (rng.random((96, 48, 64)) > 0.96).astype(np.uint8)
    k = gauss_kernel(1.5).astype(np.float32)
    reps = 5
    device_budget(None)
    q0, u0 = gpu.diffusion(m, k, reps, precision=prec, f32=True)
    # room for about 40 layers of mask, state and output: chunks of >= 16 layers
    device_budget(40 * m.shape[1] * m.shape[2] * 11)
    _say("budget fallback", prec)
    q1, u1 = gpu.diffusion(m, k, reps, precision=prec, f32=True, verbose=1)
    assert np.array_equal(u0.view(np.uint32), u1.view(np.uint32)) and np.array_equal(q0, q1)
    if prec == "fp32":
        out = np.zeros(m.shape, np.uint16)
        diffusion(m, k, out, reps)
        assert np.array_equal(out, q0)
    device_budget(None)


def _cavity_mask():
    """Solid block with small cavities: cavity voxels approach 1 and reach
    exactly 1.0f, where the reference's (u == 1.0f) rule and the true mask
    can part ways."""
    shape = (53, 24, 40)
    m = np.zeros(shape, np.uint8)
    m[8:45, 2:22, 3:37] = 1
    rng = np.random.default_rng(5)
    for _ in range(30):
        a = rng.integers(1, 5, 3)
        o = [rng.integers(10, 43 - a[0]), rng.integers(4, 20 - a[1]), rng.integers(5, 35 - a[2])]
        m[o[0]:o[0] + a[0], o[1]:o[1] + a[1], o[2]:o[2] + a[2]] = 0
    return m


def test_on_disk_reference_mask_rule(gpu, tmp_path):
    """The opt-in reference on-disk semantics: the mask is re-derived as
    (u == 1.0f) at every chunk visit (cpu:331-337, :445).  With the reference
    CPU arithmetic the uint16 file equals the oracle's quirk mode (itself
    pinned to the reference's own on-disk build) bit for bit; the MFMA form
    is within 1 LSB."""
    m = _cavity_mask()
    shape = m.shape
    k = gauss_kernel(1.1).astype(np.float32)
    fin = str(tmp_path / "mask.uint8")
    m.tofile(fin)
    for chunk_z in (9, 16, 53):
        for reps in (1, 10, 20):
            u_ref = oracle.diffusion_chunked_f32(m, k, reps, chunk_z, quirk_mask=True)
            q_ref = oracle.quantise_u16(u_ref)
            for prec in ("fp32_exact", "fp32"):
                fout = str(tmp_path / f"out_{chunk_z}_{reps}_{prec}.uint16")
                _say("reference mask rule", chunk_z, reps, prec)
                gpu.diffusion_on_disk(fin, k, fout, shape, chunk_z, reps, precision=prec, reference_mask=True)
                got = np.fromfile(fout, np.uint16).reshape(shape)
                if prec == "fp32_exact":
                    assert np.array_equal(got, q_ref)
                else:
                    assert _u16_close(got, q_ref)
    # The Gaussian case leaves the two rules within a float ulp of each other.
    # Dyadic taps with a negative centre (values stay dyadic, so every form
    # computes them exactly) make a voxel pass through exactly 1.0f and move on
    # under the true mask, while the reference's rule pins it at 1: the rules
    # then differ by whole uint16 ranges.  Compared where both results lie in
    # [0, 1] (elsewhere uint16 conversion wraps).
    taps = np.array([1.0, -0.5, 0.5], np.float32)
    for chunk_z, reps in ((9, 2), (16, 3)):
        u_ref = oracle.diffusion_chunked_f32(m, taps, reps, chunk_z, quirk_mask=True)
        u_true = oracle.diffusion_chunked_f32(m, taps, reps, chunk_z)
        ok = (u_ref >= 0) & (u_ref <= 1) & (u_true >= 0) & (u_true <= 1)
        q_ref, q_true = oracle.quantise_u16(u_ref), oracle.quantise_u16(u_true)
        assert np.count_nonzero((q_ref != q_true) & ok) > 1000
        for prec in ("fp32_exact", "fp32"):
            fout = str(tmp_path / f"dyadic_{chunk_z}_{reps}_{prec}.uint16")
            _say("reference mask rule, dyadic taps", chunk_z, reps, prec)
            gpu.diffusion_on_disk(fin, taps, fout, shape, chunk_z, reps, precision=prec, reference_mask=True)
            assert np.array_equal(np.fromfile(fout, np.uint16).reshape(shape)[ok], q_ref[ok])
            gpu.diffusion_on_disk(fin, taps, fout, shape, chunk_z, reps, precision=prec)
            assert np.array_equal(np.fromfile(fout, np.uint16).reshape(shape)[ok], q_true[ok])
    with pytest.raises(ValueError):
        gpu.diffusion_on_disk(fin, k, str(tmp_path / "x.uint16"), shape, 9, 2, precision="bf16",
                              reference_mask=True)


@pytest.mark.parametrize("prec", ["bf16", "fp32_exact"])
def test_on_disk_precisions(gpu, tmp_path, prec):
    """The file-backed form with any precision equals the in-memory form."""
    rng = np.random.default_rng(9)
    shape = (70, 20, 36)
    m = (rng.random(shape) > 0.95).astype(np.uint8)
    k = gauss_kernel(1.8).astype(np.float32)
    fin, fout = str(tmp_path / "mask.uint8"), str(tmp_path / "out.uint16")
    m.tofile(fin)
    _say("on-disk", prec)
    gpu.diffusion_on_disk(fin, k, fout, shape, 12, 5, precision=prec)
    got = np.fromfile(fout, np.uint16).reshape(shape)
    want = gpu.diffusion(m, k, 5, precision=prec)[0]
    assert np.array_equal(got, want)


# ---------------------------------------------------------------- split form
def test_split_form_takes_the_split_passes(gpu):
    """On volumes whose x extent is a multiple of 8 the split form runs its
    own passes (its result differs from the FP32 MFMA chain in the last
    bits); elsewhere it runs the FP32 form (bit-equal to it)."""
    rng = np.random.default_rng(55)
    k = gauss_kernel(5.0).astype(np.float32)
    m = (rng.random((40, 48, 64)) > 0.96).astype(np.uint8)
    a = gpu.diffusion(m, k, 4, precision="fp32_split", u16=False, f32=True)[1]
    b = gpu.diffusion(m, k, 4, precision="fp32", u16=False, f32=True)[1]
    assert not np.array_equal(a, b) and np.abs(a - b).max() <= 1e-5
    m = (rng.random((40, 48, 60)) > 0.96).astype(np.uint8)
    a = gpu.diffusion(m, k, 4, precision="fp32_split", u16=False, f32=True)[1]
    b = gpu.diffusion(m, k, 4, precision="fp32", u16=False, f32=True)[1]
    assert np.array_equal(a, b)


@pytest.mark.parametrize("sigma,reps", [(1.0, 30), (3.0, 40), (10.0, 40), (15.0, 12)])
def test_split_form_long_runs(gpu, sigma, reps):
    """Many repetitions on a workload crop: the split form stays within the
    FP32 bar of the oracle (its state carries 22 significant bits)."""
    m = WORKLOADS["c1"].mask()[:96, :104, :112].copy()
    k = gauss_kernel(sigma).astype(np.float32)
    _say("split long run", sigma, reps)
    _check_f32(gpu, m, k, reps, "fp32_split")


def test_split_form_taps_of_any_scale(gpu):
    """The taps are scaled by a power of two before the FP16 split: kernels
    with large, tiny and negative taps keep the FP32 bar."""
    rng = np.random.default_rng(12)
    m = (rng.random((24, 32, 48)) > 0.95).astype(np.uint8)
    for scale in (1e-4, 1.0, 3.0):
        k = (rng.random(21).astype(np.float32) - 0.2) * np.float32(scale)
        k /= np.float32(max(1.0, float(np.abs(k).sum())))
        _say("split taps", scale)
        u_ref, q_ref = oracle.diffusion_f32(m, k, 3)
        q, u = gpu.diffusion(m, k, 3, precision="fp32_split", u16=True, f32=True)
        assert np.abs(u - u_ref).max() <= F32_TOL


# ---------------------------------------------------------------- fused x -> y pass
FUSED_SHAPES = [(20, 70, 128), (9, 33, 136), (12, 100, 256), (7, 64, 1000), (5, 130, 72), (3, 1, 8), (4, 257, 16)]


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp32_split"])
def test_fused_pass_against_three_pass(gpu, prec):
    """The fused x -> y pass (x then y through LDS) against the three-pass
    form (y then x through HBM) and the oracle: partial strips (x extent not a
    multiple of the 128-column strip), row counts not a multiple of 32, single
    rows, every radius bucket.  The two orders round differently, so the bars
    are the form's own: 2e-3 of FP64 for the 16-bit forms, 1e-5 and 1 LSB of
    the FP32 oracle for the split form."""
    rng = np.random.default_rng(808)
    for shape in FUSED_SHAPES:
        m = (rng.random(shape) > 0.93).astype(np.uint8)
        for sigma in (0.8, 3.0, 10.0, 12.0, 16.0):
            k = gauss_kernel(sigma).astype(np.float32)
            reps = 3
            _say("fused", prec, shape, sigma)
            q, u = gpu.diffusion(m, k, reps, precision=prec, u16=True, f32=True)
            os.environ["MXB_XY_FUSED"] = "0"
            try:
                q3, u3 = gpu.diffusion(m, k, reps, precision=prec, u16=True, f32=True)
            finally:
                os.environ.pop("MXB_XY_FUSED", None)
            if prec == "fp32_split":
                u_ref, q_ref = oracle.diffusion_f32(m, k, reps)
                assert np.abs(u - u_ref).max() <= F32_TOL and _u16_close(q, q_ref)
                assert np.abs(u - u3).max() <= F32_TOL
            else:
                u64 = oracle.diffusion_f64(m, k.astype(np.float64), reps)
                assert np.abs(u - u64).max() <= LOWP_TOL
                assert np.abs(u - u3).max() <= 2 * LOWP_TOL
            assert np.all(u[m > 0] == 1.0)


def test_fused_pass_is_taken(gpu):
    """The default 16-bit path runs the fused pass where it applies: its
    rounding differs from the three-pass form's."""
    rng = np.random.default_rng(3)
    m = (rng.random((16, 64, 128)) > 0.95).astype(np.uint8)
    k = gauss_kernel(4.0).astype(np.float32)
    a = gpu.diffusion(m, k, 4, precision="bf16", u16=False, f32=True)[1]
    os.environ["MXB_XY_FUSED"] = "0"
    try:
        b = gpu.diffusion(m, k, 4, precision="bf16", u16=False, f32=True)[1]
    finally:
        os.environ.pop("MXB_XY_FUSED", None)
    assert not np.array_equal(a, b)
