// Ceiling probes: the rates the diffusion kernels are measured against.
//
// HBM: a pure read, a pure write and a copy of 16-byte vectors, each thread
// moving four vectors 4 KiB apart in its block's 16 KiB span (coalesced,
// one pass over the buffer, no grid-stride loop); the copy once more with
// non-temporal loads and stores.  MFMA: eight independent accumulator chains
// per wave of v_mfma_f32_16x16x4f32 or v_mfma_f32_16x16x32_bf16, so the
// matrix cores never wait on a dependency.  Each entry times its launches
// with HIP events on the caller's stream and returns the mean per launch.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "maxibone_diffusion.h"
#include "maxibone_probe.h"

namespace mxb {
namespace {

typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int kProbeThreads = 256;
constexpr int kProbeVecs = 4;                                   // vectors per thread
constexpr int64_t kProbeSpan = 16 * kProbeThreads * kProbeVecs;  // bytes per block

template <int KIND>
__global__ __launch_bounds__(kProbeThreads) void k_probe_hbm(const v4u *__restrict__ src, v4u *__restrict__ dst) {
    const int64_t base = (int64_t)blockIdx.x * (kProbeThreads * kProbeVecs) + threadIdx.x;
    if (KIND == MXB_PROBE_WRITE) {
        const v4u v = {(uint32_t)base, 1u, 2u, 3u};
#pragma unroll
        for (int u = 0; u < kProbeVecs; ++u) dst[base + u * kProbeThreads] = v;
        return;
    }
    v4u v[kProbeVecs];
#pragma unroll
    for (int u = 0; u < kProbeVecs; ++u)
        v[u] = KIND == MXB_PROBE_COPY_NT ? __builtin_nontemporal_load(src + base + u * kProbeThreads)
                                         : src[base + u * kProbeThreads];
    if (KIND == MXB_PROBE_READ) {
        uint32_t x = 0;
#pragma unroll
        for (int u = 0; u < kProbeVecs; ++u) x ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
        if (x == 0x9e3779b9u) dst[0] = v[0];  // keeps the loads; the buffers never hold this
        return;
    }
#pragma unroll
    for (int u = 0; u < kProbeVecs; ++u) {
        if (KIND == MXB_PROBE_COPY_NT)
            __builtin_nontemporal_store(v[u], dst + base + u * kProbeThreads);
        else
            dst[base + u * kProbeThreads] = v[u];
    }
}

constexpr int kMfmaChains = 8;
constexpr int kMfmaBlocks = 1024;  // x 4 waves: four waves per SIMD on 256 CUs

template <int KIND>
__global__ __launch_bounds__(kProbeThreads) void k_probe_mfma(int iters, float *sink) {
    const float f = 1.0f + 1e-3f * (float)(threadIdx.x & 63);
    v4f acc[kMfmaChains];
#pragma unroll
    for (int c = 0; c < kMfmaChains; ++c) acc[c] = v4f{f + (float)c, 0.f, 0.f, 0.f};
    // one scheduling barrier per trip keeps each chain in its own registers
    // (without it the 16-bit form's accumulators were rotated through copies)
    if constexpr (KIND == MXB_PROBE_MFMA_F32) {
        const float a = f, b = 1.0f - f;
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int c = 0; c < kMfmaChains; ++c) acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[c], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
    } else {
        v8bf a, b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            a[j] = (__bf16)(f + (float)j);
            b[j] = (__bf16)(f - (float)j);
        }
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int c = 0; c < kMfmaChains; ++c) acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[c], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    v4f t = acc[0];
#pragma unroll
    for (int c = 1; c < kMfmaChains; ++c) t += acc[c];
    if (t.x + t.y + t.z + t.w == 1234.5f) sink[0] = t.x;  // keeps the products; the sums never equal it
}

// Times `reps` launches of `launch` after one untimed launch.
template <class F>
int timed_launches(F launch, int reps, hipStream_t s, float *ms) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MXB_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = MXB_EDEVICE;
    if (rc == MXB_OK) {
        launch();
        bool ok = hipEventRecord(e0, s) == hipSuccess;
        for (int r = 0; r < reps; ++r) launch();
        ok = ok && hipEventRecord(e1, s) == hipSuccess;
        float t = 0.f;
        if (!ok || hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess)
            rc = MXB_EDEVICE;
        else
            *ms = t / (float)reps;
    }
    if (e0 && hipEventDestroy(e0) != hipSuccess) rc = MXB_EDEVICE;
    if (e1 && hipEventDestroy(e1) != hipSuccess) rc = MXB_EDEVICE;
    return rc;
}

}  // namespace
}  // namespace mxb

extern "C" int mxb_probe_hbm(int kind, const void *src, void *dst, int64_t bytes, int reps, void *stream,
                             float *ms_per_launch) {
    using namespace mxb;
    if (ms_per_launch == nullptr || reps < 1 || bytes <= 0 || bytes % kProbeSpan != 0) return MXB_EINVAL;
    if (kind < MXB_PROBE_READ || kind > MXB_PROBE_COPY_NT) return MXB_EINVAL;
    if (kind != MXB_PROBE_WRITE && src == nullptr) return MXB_EINVAL;
    if (dst == nullptr || (reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) % 16 != 0)
        return MXB_EINVAL;
    const int64_t blocks = bytes / kProbeSpan;
    if (blocks > 0x7fffffff) return MXB_EINVAL;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const v4u *in = static_cast<const v4u *>(src);
    v4u *out = static_cast<v4u *>(dst);
    auto launch = [&]() {
        const dim3 g((unsigned)blocks), b(kProbeThreads);
        switch (kind) {
            case MXB_PROBE_READ: hipLaunchKernelGGL(k_probe_hbm<MXB_PROBE_READ>, g, b, 0, s, in, out); break;
            case MXB_PROBE_WRITE: hipLaunchKernelGGL(k_probe_hbm<MXB_PROBE_WRITE>, g, b, 0, s, in, out); break;
            case MXB_PROBE_COPY: hipLaunchKernelGGL(k_probe_hbm<MXB_PROBE_COPY>, g, b, 0, s, in, out); break;
            default: hipLaunchKernelGGL(k_probe_hbm<MXB_PROBE_COPY_NT>, g, b, 0, s, in, out); break;
        }
    };
    return timed_launches(launch, reps, s, ms_per_launch);
}

extern "C" int mxb_probe_mfma(int kind, int64_t iters, void *sink, int reps, void *stream, float *ms_per_launch,
                              double *flops_per_launch) {
    using namespace mxb;
    if (ms_per_launch == nullptr || flops_per_launch == nullptr || sink == nullptr || iters < 1 ||
        iters > (1 << 24) || reps < 1)
        return MXB_EINVAL;
    if (kind != MXB_PROBE_MFMA_F32 && kind != MXB_PROBE_MFMA_BF16) return MXB_EINVAL;
    hipStream_t s = static_cast<hipStream_t>(stream);
    float *out = static_cast<float *>(sink);
    auto launch = [&]() {
        if (kind == MXB_PROBE_MFMA_F32)
            hipLaunchKernelGGL(k_probe_mfma<MXB_PROBE_MFMA_F32>, dim3(kMfmaBlocks), dim3(kProbeThreads), 0, s, (int)iters, out);
        else
            hipLaunchKernelGGL(k_probe_mfma<MXB_PROBE_MFMA_BF16>, dim3(kMfmaBlocks), dim3(kProbeThreads), 0, s, (int)iters, out);
    };
    // 2 M N K flops per MFMA: 16x16x4 -> 2048, 16x16x32 -> 16384
    const double per = kind == MXB_PROBE_MFMA_F32 ? 2048.0 : 16384.0;
    *flops_per_launch = per * kMfmaChains * (double)iters * (kMfmaBlocks * (kProbeThreads / 64));
    return timed_launches(launch, reps, s, ms_per_launch);
}
