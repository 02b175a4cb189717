"""mxb_device_diffusion_noalloc: the variant of mxb_device_diffusion whose
state buffers live in the library's fixed per-device array.

Checked through the C ABI against mxb_device_diffusion (same kernels, so the
bar is bit equality), on two calls in a row on one stream, the second with a
smaller volume, so that contents the first call left in the shared array
cannot leak into the second.  Volumes beyond the array fail with EINVAL.
"""
import sys

import numpy as np
import pytest

from maxibone_amd.helpers import gauss_kernel

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp32_exact"])
def test_noalloc_equals_device_diffusion(gpu, prec):
    from maxibone_amd import hip
    L = gpu.lib()
    P = gpu.PRECISION[prec]
    rng = np.random.default_rng(7)
    # the second volume is smaller; 128^3 is the documented size limit
    cases = [((128, 128, 128), 10.0, 4), ((40, 50, 72), 2.5, 3)]
    st = hip.Stream()
    for shape, sigma, reps in cases:
        _say("noalloc", prec, shape, sigma, reps)
        mask = (rng.random(shape) > 0.97).astype(np.uint8)
        k = gauss_kernel(sigma).astype(np.float32)
        n = mask.size
        sb = L.mxb_state_bytes(P, *shape)
        assert ((sb + 255) // 256 * 256 + sb) <= gpu.NOALLOC_SCRATCH_BYTES
        dm = hip.DeviceBuffer.from_array(mask)
        s0, s1 = hip.DeviceBuffer(sb), hip.DeviceBuffer(sb)
        want_q, want_f = hip.DeviceBuffer(2 * n), hip.DeviceBuffer(4 * n)
        got_q, got_f = hip.DeviceBuffer(2 * n), hip.DeviceBuffer(4 * n)
        for b in (got_q, got_f):
            b.fill_bytes(0xFF)
        gpu.check(L.mxb_device_diffusion(dm.ptr, *shape, k.ctypes.data, len(k), reps, P, s0.ptr, s1.ptr,
                                          want_q.ptr, want_f.ptr, st.handle))
        gpu.check(L.mxb_device_diffusion_noalloc(dm.ptr, *shape, k.ctypes.data, len(k), reps, P,
                                                 got_q.ptr, got_f.ptr, st.handle))
        st.synchronize()
        wf, gf = want_f.to_array(np.float32, shape), got_f.to_array(np.float32, shape)
        wq, gq = want_q.to_array(np.uint16, shape), got_q.to_array(np.uint16, shape)
        assert np.array_equal(gf.view(np.uint32), wf.view(np.uint32))
        assert np.array_equal(gq, wq)
        # uint16 output alone
        got_q.fill_bytes(0)
        gpu.check(L.mxb_device_diffusion_noalloc(dm.ptr, *shape, k.ctypes.data, len(k), reps, P,
                                                 got_q.ptr, None, st.handle))
        st.synchronize()
        assert np.array_equal(got_q.to_array(np.uint16, shape), wq)
        for b in (dm, s0, s1, want_q, want_f, got_q, got_f):
            b.free()


def test_noalloc_rejects_volumes_beyond_the_array(gpu):
    from maxibone_amd import hip
    L = gpu.lib()
    shape = (160, 128, 128)  # 2 x 10.7 MB of state
    k = gauss_kernel(2.0).astype(np.float32)
    dm = hip.DeviceBuffer(int(np.prod(shape)))
    out = hip.DeviceBuffer(2 * int(np.prod(shape)))
    with pytest.raises(ValueError):
        gpu.check(L.mxb_device_diffusion_noalloc(dm.ptr, *shape, k.ctypes.data, len(k), 2,
                                                 gpu.PRECISION["fp32"], out.ptr, None, None))
    dm.free()
    out.free()
