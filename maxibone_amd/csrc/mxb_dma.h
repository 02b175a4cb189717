// Global -> LDS direct loads (gfx950 global_load_lds_dwordx4) with explicit
// wait counts.
//
// The loads are issued from inline assembly so that the compiler's wait-count
// tracking does not see them: it would otherwise wait for every outstanding
// LDS-bound load before any LDS read it cannot prove disjoint, which drains a
// multi-slot prefetch ring on every iteration.  Callers therefore wait
// themselves, with vm_wait<N>(): "at most N vector-memory operations of this
// wave still outstanding".  Vector-memory operations complete in issue order,
// so N must be at most the number of vector-memory instructions the wave
// issued after the load being waited for; kernels built on this keep that
// number a compile-time constant (no conditional loads or stores in the
// pipelined part of their loops, no register spills).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mxb {

typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));
typedef float v4f32 __attribute__((ext_vector_type(4)));

// 16-byte global store; NT: non-temporal (streaming) policy.  Measured on
// MI355X: it helps passes that write whole 256-byte rows and hurts those that
// write 64-128-byte pieces, which rely on L2 write combining.  The pointer is
// taken as global memory explicitly: where it comes from a descriptor in
// memory the compiler could not infer that and would emit a flat store, which
// also counts against the LDS wait counter.
typedef __attribute__((address_space(1))) v4u32 g_v4u32;
typedef __attribute__((address_space(1))) v4f32 g_v4f32;

// `p` as a global-memory pointer (see store16).
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem(T *p) {
    return (__attribute__((address_space(1))) T *)p;
}

#ifndef MXB_XCD_SWIZZLE
#define MXB_XCD_SWIZZLE 1
#endif
#ifndef MXB_NT_ALIGNED_ONLY
#define MXB_NT_ALIGNED_ONLY 1
#endif

// Block b of a b_count-block grid is dispatched to XCD b % 8.  xcd_tile(b)
// numbers the tiles so that each XCD gets one contiguous range of them:
// neighbouring tiles, whose rows share L2 lines when a row is not 128-byte
// aligned, then run on the same XCD at about the same time.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t b_count) {
    if (!MXB_XCD_SWIZZLE) return b;
    const int64_t q = b_count / 8, rem = b_count % 8, x = b % 8;
    return x * q + (x < rem ? x : rem) + b / 8;
}

// Whether rows of `row_bytes` from `base` may use the non-temporal store
// policy: only whole 128-byte lines bypass L2 write combining safely.
__device__ __forceinline__ bool nt_rows_ok(const void *base, int64_t row_bytes) {
    if (!MXB_NT_ALIGNED_ONLY) return true;
    return row_bytes % 128 == 0 && reinterpret_cast<uintptr_t>(base) % 128 == 0;
}

template <bool NT>
__device__ __forceinline__ void store16(void *p, uint4 v) {
    const v4u32 x = {v.x, v.y, v.z, v.w};
    g_v4u32 *q = (g_v4u32 *)p;
    if (NT)
        __builtin_nontemporal_store(x, q);
    else
        *q = x;
}
template <bool NT>
__device__ __forceinline__ void store16(float *p, v4f32 v) {
    g_v4f32 *q = (g_v4f32 *)p;
    if (NT)
        __builtin_nontemporal_store(v, q);
    else
        *q = v;
}

// LDS byte address of a __shared__ object.
__device__ __forceinline__ uint32_t lds_addr(const void *p) {
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p;
}

// 16 bytes per active lane: lane i writes LDS bytes [base + 16 i, +16) from
// the 16-byte-aligned global address `src`.  `lds_wave_base` must be uniform.
// NT: non-temporal policy (same trade-off as store16).
template <bool NT>
__device__ __forceinline__ void dma16(uint32_t lds_wave_base, const void *src) {
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
    // one wait state between the SALU write of M0 and the LDS-DMA read of it
    if (NT)
        asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off nt"
                     :
                     : "s"(lds_wave_base), "v"(src)
                     : "memory", "m0");
    else
        asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                     :
                     : "s"(lds_wave_base), "v"(src)
                     : "memory", "m0");
#pragma clang diagnostic pop
}

// 4 bytes per active lane: lane i writes LDS bytes [base + 4 i, +4) from the
// 4-byte-aligned global address `src` (same wait-count rules as dma16).
__device__ __forceinline__ void dma4(uint32_t lds_wave_base, const void *src) {
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dword %1, off"
                 :
                 : "s"(lds_wave_base), "v"(src)
                 : "memory", "m0");
#pragma clang diagnostic pop
}

// s_waitcnt with vmcnt = N and no wait on the other counters.
template <int N>
__device__ __forceinline__ void vm_wait() {
    static_assert(N >= 0 && N < 64, "vmcnt is a 6-bit field");
    __builtin_amdgcn_s_waitcnt((N & 0xf) | ((N >> 4) << 14) | (0x7 << 4) | (0xf << 8));
}

// Zero bytes: the source of loads that fall outside the volume, so every
// lane of every wave issues the same number of loads.  (One copy per
// translation unit; device globals start zeroed and are never written.)
static __device__ __attribute__((aligned(16))) uint4 g_dma_zero[4];
// Destination of the stores of lanes whose line lies beyond the volume (the
// store is still issued, keeping the per-iteration count fixed).
static __device__ __attribute__((aligned(16))) uint4 g_dma_sink[64];

}  // namespace mxb
