"""The diffusion workloads at their real sizes (needs a GPU; the
2048x2048x1536 volume needs ~100 GB of device memory and ~110 GB of host
memory).

* 1024^3: cylinder r=300 + 2000 spheres, sigma=10 (R=40), 40 repetitions.
* 2048x2048x1536: cylinder r=700 + 20000 spheres, sigma=15 (R=60), 100
  repetitions; in HBM and through the z-chunked path (chunks of 256 layers,
  halo = R, the reference's one-step-per-chunk scheme).
* 64 sub-volumes, dims U[64,384], sigma~U[2,10], 20 repetitions (batched).

Window checks: the GPU diffuses the whole volume for T = 2 repetitions, and
32^3 windows of the result are compared with the oracle run on the window
plus a T*R halo on every side (clipped at the volume faces).  That crop's
interior is exact: the crop's own zero boundary differs from the volume only
on faces that are not volume faces, and each repetition carries the error at
most R voxels further along each axis, so after T repetitions every voxel at
least T*R from such a face has the full volume's value, computed by the same
arithmetic in the same order.  Windows sit at corners, on faces, across the
cylinder surface and, for 2048x2048x1536, at flat offsets beyond 2^31 and
2^32 voxels.  Bars (src/test/test_diffusion.py:92-98 for FP32):
  fp32_exact: bit-equal to the oracle;
  fp32 (MFMA) and fp32_split (FP16 hi + lo state, three FP16 MFMAs per
  chunk): max-abs <= 1e-5 and uint16 within 1 LSB of the oracle;
  bf16: max-abs <= 2e-3 of the FP64 oracle.

Full-repetition runs check what exact arithmetic fixes: u = 1 on the mask,
0 <= u <= 1, and u non-decreasing from one repetition to the next (G >= 0
and u1 >= u0), with a slack for rounding.
"""
import sys

import numpy as np
import pytest

import oracle
from maxibone_amd.synth import WORKLOADS, batched_volumes

pytestmark = pytest.mark.gpu
F32_TOL = 1e-5
LOWP_TOL = 2e-3
T = 2
W = 32


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oracle_window(mask, k, reps, origin, f64=False):
    """The oracle's result on the W^3 window at `origin` (f32: (u, q); f64: u)."""
    R = len(k) // 2
    H = reps * R
    lo = [max(0, o - H) for o in origin]
    hi = [min(n, o + W + H) for o, n in zip(origin, mask.shape)]
    crop = np.ascontiguousarray(mask[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]])
    sl = tuple(slice(o - l, o - l + W) for o, l in zip(origin, lo))
    if f64:
        return oracle.diffusion_f64(crop, k.astype(np.float64), reps)[sl]
    u, q = oracle.diffusion_f32(crop, k, reps)
    return u[sl], q[sl]


class DeviceRun:
    """One workload volume resident in HBM: mask, two state buffers and both
    outputs, diffused through mxb_device_diffusion."""

    def __init__(self, capi, hip, mask, k):
        self.capi, self.hip = capi, hip
        self.shape = mask.shape
        self.k = np.ascontiguousarray(k, np.float32)
        n = mask.size
        L = capi.lib()
        sb = L.mxb_state_bytes(capi.PRECISION["fp32"], *self.shape)
        self.dm = hip.DeviceBuffer.from_array(mask)
        self.s0, self.s1 = hip.DeviceBuffer(sb), hip.DeviceBuffer(sb)
        self.dq, self.df = hip.DeviceBuffer(2 * n), hip.DeviceBuffer(4 * n)

    def run(self, reps, precision, f32=True):
        L = self.capi.lib()
        self.capi.check(L.mxb_device_diffusion(self.dm.ptr, *self.shape, self.k.ctypes.data, len(self.k),
                                               int(reps), self.capi.PRECISION[precision], self.s0.ptr,
                                               self.s1.ptr, self.dq.ptr, self.df.ptr if f32 else None, None))
        self.hip.synchronize()

    def window_f32(self, origin):
        return self.df.box(np.float32, self.shape, origin, (W, W, W))

    def window_u16(self, origin):
        return self.dq.box(np.uint16, self.shape, origin, (W, W, W))

    def free(self):
        for b in (self.dm, self.s0, self.s1, self.dq, self.df):
            b.free()


def _check_windows(dev, mask, k, origins, precisions):
    for prec in precisions:
        dev.run(T, prec)
        for o in origins:
            flat = (o[0] * mask.shape[1] + o[1]) * mask.shape[2] + o[2]
            u, q = dev.window_f32(o), dev.window_u16(o)
            if prec == "bf16":
                u64 = _oracle_window(mask, k, T, o, f64=True)
                err = float(np.abs(u - u64).max())
                _say(f"window {o} (flat offset {flat}) {prec}: max-abs vs FP64 {err:.2e}")
                assert err <= LOWP_TOL
                continue
            u_ref, q_ref = _oracle_window(mask, k, T, o)
            err = float(np.abs(u - u_ref).max())
            dq = int(np.abs(q.astype(np.int32) - q_ref.astype(np.int32)).max())
            _say(f"window {o} (flat offset {flat}) {prec}: max-abs {err:.2e}, uint16 {dq} LSB")
            if prec == "fp32_exact":
                assert np.array_equal(u.view(np.uint32), u_ref.view(np.uint32))
                assert np.array_equal(q, q_ref)
            else:
                assert err <= F32_TOL and dq <= 1


# Slack of the monotonicity check: FP32 rounds each pass to ~1 ulp, so where
# the state has converged a repetition may lower it by a few ulps; the 16-bit
# forms round the state to BF16 between repetitions, which can move it by up
# to their error bar.
MONO_SLACK = {"fp32": lambda u: 5e-7 * u, "fp32_split": lambda u: 5e-7 * u, "bf16": lambda u: LOWP_TOL}


def _check_properties(u_prev, u, mask_slab, prec):
    """u = 1 on the mask, 0 <= u <= 1, u >= u_prev (up to rounding)."""
    on = mask_slab > 0
    assert np.all(u[on] == 1.0)
    assert float(u.min()) >= 0.0 and float(u.max()) <= 1.0 + 1e-6
    drop = float((u_prev - u - MONO_SLACK[prec](u_prev)).max())
    assert drop <= 0.0, f"state decreased by {drop} beyond the rounding slack"


# ------------------------------------------------------------------ 1024^3
C2_ORIGINS = [(0, 0, 0), (992, 992, 992), (0, 496, 496), (496, 0, 496), (496, 496, 992),
              (496, 496, 796), (300, 500, 500), (717, 123, 345)]


@pytest.fixture(scope="module")
def c2(gpu, workload_mask):
    from maxibone_amd import hip
    w = WORKLOADS["c2"]
    mask = workload_mask("c2")
    k = w.kernel().astype(np.float32)
    dev = DeviceRun(gpu, hip, mask, k)
    yield w, mask, k, dev
    dev.free()


def test_c2_windows(c2):
    w, mask, k, dev = c2
    _check_windows(dev, mask, k, C2_ORIGINS, ("fp32_exact", "fp32", "fp32_split", "bf16"))


def test_c2_full_repetitions_properties(c2):
    """All 40 repetitions in FP32 and BF16: the invariants on the whole volume
    (repetitions 39 and 40 compared slab by slab)."""
    w, mask, k, dev = c2
    nz = mask.shape[0]
    for prec in ("fp32", "fp32_split", "bf16"):
        _say("1024^3 full repetitions", prec)
        dev.run(w.reps - 1, prec)
        prev = dev.df.to_array(np.float32, mask.shape)
        dev.run(w.reps, prec)
        for z0 in range(0, nz, 128):
            u = dev.df.box(np.float32, mask.shape, (z0, 0, 0), (min(128, nz - z0),) + mask.shape[1:])
            _check_properties(prev[z0:z0 + u.shape[0]], u, mask[z0:z0 + u.shape[0]], prec)
        # uint16 output is the quantised float state (cpu:191-196)
        q = dev.dq.to_array(np.uint16, mask.shape)
        for z0 in range(0, nz, 128):
            u = dev.df.box(np.float32, mask.shape, (z0, 0, 0), (min(128, nz - z0),) + mask.shape[1:])
            assert np.array_equal(q[z0:z0 + u.shape[0]], oracle.quantise_u16(u))
        del prev, q


# ------------------------------------------------------------------ 2048x2048x1536
# flat offsets: z >= 683 passes 2^31 voxels, z >= 1366 passes 2^32
C4_ORIGINS = [(0, 0, 0), (2016, 2016, 1504), (700, 1008, 1452), (1400, 1008, 0), (1400, 0, 752),
              (2016, 1008, 752), (1024, 1708, 752), (1100, 300, 200)]


@pytest.fixture(scope="module")
def c4(gpu, workload_mask):
    from maxibone_amd import hip
    w = WORKLOADS["c4"]
    mask = workload_mask("c4")
    k = w.kernel().astype(np.float32)
    dev = DeviceRun(gpu, hip, mask, k)
    yield w, mask, k, dev
    dev.free()


@pytest.mark.timeout(600)
def test_c4_windows(c4):
    w, mask, k, dev = c4
    _check_windows(dev, mask, k, C4_ORIGINS, ("fp32_exact", "fp32", "fp32_split"))
    _check_windows(dev, mask, k, C4_ORIGINS[:5], ("bf16",))


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec", ["fp32", "fp32_split", "bf16"])
def test_c4_chunked_equals_in_core(c4, gpu, prec):
    """z-chunks of 256 layers with halo R (one repetition per chunk visit,
    state in host memory between visits): bit-equal to the in-HBM result."""
    w, mask, k, dev = c4
    _say("2048x2048x1536 chunked z=256", prec)
    q1, u1 = gpu.diffusion_chunked(mask, k, T, 256, 1, precision=prec, u16=True, f32=True)
    dev.run(T, prec)
    nz = mask.shape[0]
    for z0 in range(0, nz, 128):
        u = dev.df.box(np.float32, mask.shape, (z0, 0, 0), (128,) + mask.shape[1:])
        q = dev.dq.box(np.uint16, mask.shape, (z0, 0, 0), (128,) + mask.shape[1:])
        assert np.array_equal(u.view(np.uint32), u1[z0:z0 + 128].view(np.uint32)), f"z0={z0}"
        assert np.array_equal(q, q1[z0:z0 + 128]), f"z0={z0}"
    del q1, u1


@pytest.mark.timeout(600)
def test_c4_full_repetitions_properties(c4):
    """All 100 repetitions, FP32 and BF16: the invariants on the whole volume,
    on the uint16 result (monotone, 65535 on the mask) and on the float
    state of a slab per 256 layers."""
    w, mask, k, dev = c4
    nz = mask.shape[0]
    for prec in ("fp32", "fp32_split", "bf16"):
        _say("2048x2048x1536 full repetitions", prec)
        dev.run(w.reps - 1, prec)
        prev_q = dev.dq.to_array(np.uint16, mask.shape)
        prev_u = [dev.df.box(np.float32, mask.shape, (z0, 0, 0), (16,) + mask.shape[1:])
                  for z0 in range(0, nz, 256)]
        dev.run(w.reps, prec)
        for i, z0 in enumerate(range(0, nz, 256)):
            u = dev.df.box(np.float32, mask.shape, (z0, 0, 0), (16,) + mask.shape[1:])
            _check_properties(prev_u[i], u, mask[z0:z0 + 16], prec)
        for z0 in range(0, nz, 128):
            q = dev.dq.box(np.uint16, mask.shape, (z0, 0, 0), (128,) + mask.shape[1:])
            on = mask[z0:z0 + 128] > 0
            assert np.all(q[on] == 65535)
            lsb = int(LOWP_TOL * 65535) + 1 if prec == "bf16" else 1
            assert np.all(q.astype(np.int32) >= prev_q[z0:z0 + 128].astype(np.int32) - lsb)
        del prev_q


# ------------------------------------------------------------------ 64 sub-volumes
@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_batch_of_64_volumes(gpu, prec):
    """The real batch (20 repetitions, autotuned plan, captured graph): every
    volume bit-equal to per-volume launches of the single-volume path; four
    small volumes (radius buckets 12, 32 and 40) against the oracle."""
    from maxibone_amd import hip
    reps = 20
    vols = batched_volumes()
    kernels = [v.kernel().astype(np.float32) for v in vols]
    dm = [hip.DeviceBuffer.from_array(v.mask) for v in vols]
    of = [hip.DeviceBuffer(4 * v.mask.size) for v in vols]
    oq = [hip.DeviceBuffer(2 * v.mask.size) for v in vols]
    _say("64-volume batch", prec)
    plan = gpu.BatchPlan([b.ptr for b in dm], [v.mask.shape for v in vols], kernels, reps,
                         outs_u16=[b.ptr for b in oq], outs_f32=[b.ptr for b in of], precision=prec)
    info = plan.info()
    _say("plan", info)
    assert info["persistent"]
    plan.run(1)
    plan.sync()
    single = [(b.to_array(np.float32, v.mask.shape), bq.to_array(np.uint16, v.mask.shape))
              for b, bq, v in zip(of, oq, vols)]
    for b in of + oq:
        b.fill_bytes(0xFF)
    hip.synchronize()
    plan.run(0)
    plan.sync()
    for i, (b, bq, v) in enumerate(zip(of, oq, vols)):
        u = b.to_array(np.float32, v.mask.shape)
        q = bq.to_array(np.uint16, v.mask.shape)
        assert np.array_equal(u.view(np.uint32), single[i][0].view(np.uint32)), f"volume {i}"
        assert np.array_equal(q, single[i][1]), f"volume {i}"
    for i in (8, 59, 60, 25):  # radii 27, 36, 28, 10: buckets 32, 40, 32, 12
        v = vols[i]
        u = single[i][0]
        if prec == "fp32":
            u_ref, q_ref = oracle.diffusion_f32(v.mask, kernels[i], reps)
            err = float(np.abs(u - u_ref).max())
            dq = int(np.abs(single[i][1].astype(np.int32) - q_ref.astype(np.int32)).max())
            _say(f"volume {i} {v.mask.shape} R={v.radius}: max-abs {err:.2e}, uint16 {dq} LSB")
            assert err <= F32_TOL and dq <= 1
        else:
            u64 = oracle.diffusion_f64(v.mask, kernels[i].astype(np.float64), reps)
            err = float(np.abs(u - u64).max())
            _say(f"volume {i} {v.mask.shape} R={v.radius}: max-abs vs FP64 {err:.2e}")
            assert err <= LOWP_TOL
    plan.close()
    for b in dm + of + oq:
        b.free()
