// Reduced-precision (BF16 / FP16 state) line passes of the masked 3D diffusion:
// register-ring form, the fallback for inputs whose rows are not 16-byte
// aligned (the LDS-staged form is in mxb_kernels_lowp_lds.hip; the arithmetic,
// chunking and tap split are shared, mxb_lowp.h).
//
// Same pass structure and data layouts as the FP32 kernels (lines grouped 64
// to a block, the y-pass writing (z, x, y) order, the x-pass illuminating).
// The inputs a lane needs are 8-position blocks; block b (positions
// a0 - R + 8b .. +7) is used by lanes with g = b mod 2 ... so lane g keeps,
// in ring slot m, block 2m + g: tile t, chunk c then reads slot t + 2c,
// which is the same register in every lane.
#include "mxb_lowp.h"

namespace mxb {

__device__ uint16_t g_zero_line16[128];

template <int RT>
struct LowpShape {
    static constexpr int C = LowpChunks<RT>::C;        // 32-position chunks per tile
    static constexpr int NT = 2;                        // tiles per group
    static constexpr int CR = 2 * C + 2;                // ring slots (one group of prefetch)
    static constexpr int P = CR / NT;                   // groups per ring period
};

struct BlockPosL {
    int64_t outer, col0, a0, a1;
};

__device__ __forceinline__ BlockPosL decode_lowp(const LineGeom &g) {
    int64_t bid = blockIdx.x;
    const int64_t seg_id = bid % g.nseg;
    bid /= g.nseg;
    const int64_t nlb = (g.n_lane + 63) >> 6;
    const int64_t lb = bid % nlb;
    bid /= nlb;
    BlockPosL p;
    p.outer = bid;
    p.col0 = lb * 64 + 16 * __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    p.a0 = seg_id * g.seg;
    p.a1 = p.a0 + g.seg < g.n_ax ? p.a0 + g.seg : g.n_ax;
    return p;
}

template <int RT, bool BF>
__global__ __launch_bounds__(256) void k_line_lowp(const uint16_t *__restrict__ in,
                                                   uint16_t *__restrict__ out,
                                                   const uint8_t *__restrict__ mask, LineGeom g,
                                                   TapsL<2 * RT + 1> taps) {
    using S = LowpShape<RT>;
    using L = Lowp<BF>;
    constexpr int C = S::C, NT = S::NT, CR = S::CR, P = S::P;
    constexpr int TW = LowpChunks<RT>::TW;
    __shared__ uint16_t thi[TW], tlo[TW];
    const BlockPosL bp = decode_lowp(g);
    const int lane = threadIdx.x & 63;
    const int col = lane & 15, kg = lane >> 4;
    lowp_fill_taps<RT>(thi, tlo, taps);
    __syncthreads();
    v8s ahi[C], alo[C];
    lowp_read_taps<RT>(thi, tlo, col, kg, ahi, alo);

    const int64_t n_ax = g.n_ax;
    const int64_t in_ax_b = g.in_ax * 2;
    const int64_t li = bp.col0 + col;
    const bool act = li < g.n_lane;
    const int64_t lc = act ? li : g.n_lane - 1;
    const char *wbase = reinterpret_cast<const char *>(in + bp.outer * g.in_outer);  // uniform
    const uint32_t lcol_b = (uint32_t)(lc * 2);
    const uint32_t loff = lcol_b + (uint32_t)(8 * kg * in_ax_b);
    const char *zsrc = reinterpret_cast<const char *>(g_zero_line16) + 2 * lane;
    const int64_t q_first = bp.a0 - RT;

    v8s ring[CR];
    // slot m holds, for lane kg, positions q_first + 16m + 8kg + e, e = 0..7
    auto load_slot = [&](int slot, int64_t m) {
        const int64_t qb = q_first + 16 * m;  // uniform
        const bool inside = qb >= 0 && qb + 31 < n_ax;
        uint16_t v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const char *addr = wbase + (qb + e) * in_ax_b + loff;
            if (!inside) {
                const int64_t p = qb + 8 * kg + e;
                addr = (p >= 0 && p < n_ax) ? addr : zsrc;
            }
            v[e] = *reinterpret_cast<const uint16_t *>(addr);
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) ring[slot][e] = (short)v[e];
    };
#pragma unroll
    for (int m = 0; m < CR; ++m) load_slot(m, m);

    const int64_t obase = bp.outer * g.out_outer + lc * g.out_lane + (bp.a0 + 4 * kg) * g.out_ax;
    const int64_t ntiles = (bp.a1 - bp.a0 + 15) / 16;
    const int64_t ngroups = (ntiles + NT - 1) / NT;
    int64_t gi = 0;
    while (true) {
#pragma unroll
        for (int j = 0; j < P; ++j) {
            v4f acc[NT];
#pragma unroll
            for (int u = 0; u < NT; ++u) acc[u] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int c = 0; c < C; ++c)
#pragma unroll
                for (int u = 0; u < NT; ++u) {
                    const v8s x = ring[(NT * j + u + 2 * c) % CR];
                    acc[u] = L::mfma(ahi[c], x, acc[u]);
                    acc[u] = L::mfma(alo[c], x, acc[u]);
                }
            // prefetch the two new slots of group gi + 2 into the slots of group gi
#pragma unroll
            for (int q = 0; q < NT; ++q) load_slot((NT * j + q) % CR, NT * gi + CR + q);
#pragma unroll
            for (int u = 0; u < NT; ++u) {
                const int64_t t = NT * gi + u;
                const int64_t tpos = bp.a0 + 16 * t;
                if (tpos >= bp.a1) break;
                if (!act) continue;
                const bool full = tpos + 16 <= bp.a1;
                float v[4] = {acc[u][0], acc[u][1], acc[u][2], acc[u][3]};
                if (full && g.vec_out) {
                    const int64_t off = obase + 16 * t;
                    if (mask != nullptr) {
                        const uint32_t mm = *reinterpret_cast<const uint32_t *>(mask + off);
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if ((mm >> (8 * r)) & 0xffu) v[r] = 1.0f;
                    }
                    *reinterpret_cast<uint2 *>(out + off) =
                        make_uint2(pack2(L::from_f32(v[0]), L::from_f32(v[1])),
                                   pack2(L::from_f32(v[2]), L::from_f32(v[3])));
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (tpos + 4 * kg + r < bp.a1) {
                            const int64_t off = obase + (16 * t + r) * g.out_ax;
                            float w = v[r];
                            if (mask != nullptr && mask[off] != 0) w = 1.0f;
                            out[off] = L::from_f32(w);
                        }
                }
            }
            if (++gi == ngroups) return;
        }
    }
}

// The element-wise kernels below take 8 elements per thread (an 8-byte mask
// load, 16-byte stores) when `vec` says the pointers allow it, and finish the
// tail, or everything when they do not, one element per thread.

// The 16-bit values (v0, v1) of a byte pair of w at bit `sh`: `one` where the
// byte is nonzero, else 0.
__device__ __forceinline__ uint32_t mask_pair16(uint32_t w, int sh, uint16_t one) {
    return (((w >> sh) & 0xffu) ? (uint32_t)one : 0u) | ((((w >> (sh + 8)) & 0xffu) ? (uint32_t)one : 0u) << 16);
}

__device__ __forceinline__ uint4 mask8_to16(uint2 mv, uint16_t one) {
    return make_uint4(mask_pair16(mv.x, 0, one), mask_pair16(mv.x, 16, one), mask_pair16(mv.y, 0, one),
                      mask_pair16(mv.y, 16, one));
}

template <bool BF>
__global__ __launch_bounds__(256) void k_init_lowp(const uint8_t *__restrict__ m,
                                                   uint16_t *__restrict__ u, int64_t n, int vec) {
    const uint16_t one = Lowp<BF>::from_f32(1.0f);
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t done = 0;
    if (vec) {
        const int64_t n8 = n >> 3;
        for (int64_t g = tid; g < n8; g += stride)
            *reinterpret_cast<uint4 *>(u + 8 * g) = mask8_to16(*reinterpret_cast<const uint2 *>(m + 8 * g), one);
        done = n8 << 3;
    }
    for (int64_t i = done + tid; i < n; i += stride) u[i] = m[i] > 0 ? one : (uint16_t)0;
}

template <bool BF>
__global__ __launch_bounds__(256) void k_lowp_to_f32(const uint16_t *__restrict__ u,
                                                     float *__restrict__ f, int64_t n, int vec) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t done = 0;
    if (vec) {
        const int64_t n8 = n >> 3;
        for (int64_t g = tid; g < n8; g += stride) {
            const uint4 v = *reinterpret_cast<const uint4 *>(u + 8 * g);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
            float r[8];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                r[2 * k] = Lowp<BF>::to_f32((uint16_t)(w[k] & 0xffffu));
                r[2 * k + 1] = Lowp<BF>::to_f32((uint16_t)(w[k] >> 16));
            }
            float4 *dst = reinterpret_cast<float4 *>(f + 8 * g);
            dst[0] = make_float4(r[0], r[1], r[2], r[3]);
            dst[1] = make_float4(r[4], r[5], r[6], r[7]);
        }
        done = n8 << 3;
    }
    for (int64_t i = done + tid; i < n; i += stride) f[i] = Lowp<BF>::to_f32(u[i]);
}

// Split-form planes from the mask: hi = f16(1) on the mask, else 0; lo = 0.
__global__ __launch_bounds__(256) void k_init_split(const uint8_t *__restrict__ m, uint16_t *__restrict__ p,
                                                    int64_t lo_off, int64_t n, int vec) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t done = 0;
    if (vec) {
        const int64_t n8 = n >> 3;
        for (int64_t g = tid; g < n8; g += stride) {
            *reinterpret_cast<uint4 *>(p + 8 * g) = mask8_to16(*reinterpret_cast<const uint2 *>(m + 8 * g), 0x3c00);
            *reinterpret_cast<uint4 *>(p + lo_off + 8 * g) = make_uint4(0u, 0u, 0u, 0u);
        }
        done = n8 << 3;
    }
    for (int64_t i = done + tid; i < n; i += stride) {
        p[i] = m[i] > 0 ? (uint16_t)0x3c00 : (uint16_t)0;
        p[lo_off + i] = 0;
    }
}

// Split-form planes from an FP32 state.
__global__ __launch_bounds__(256) void k_split_state(const float *__restrict__ u, uint16_t *__restrict__ p,
                                                     int64_t lo_off, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint16_t hi, lo;
        split_f16(u[i], hi, lo);
        p[i] = hi;
        p[lo_off + i] = lo;
    }
}

// ---------------------------------------------------------------- host side
static void plan_segments_lowp(LineGeom &g) {
    const int64_t threads = ((g.n_lane + 63) / 64) * 16 * g.n_outer;  // one wave per 16 lines
    const int64_t target = 256LL * 16 * 16;
    int64_t nseg = threads > 0 ? (target + threads - 1) / threads : 1;
    const int64_t max_seg = (g.n_ax + 63) / 64;
    if (nseg > max_seg) nseg = max_seg;
    if (nseg < 1) nseg = 1;
    int64_t seg = (g.n_ax + nseg - 1) / nseg;
    seg = ((seg + 31) / 32) * 32;
    g.seg = seg;
    g.nseg = (g.n_ax + seg - 1) / seg;
    if (g.nseg < 1) g.nseg = 1;
}

template <int RT>
static hipError_t launch_lowp_rt(const uint16_t *in, uint16_t *out, const uint8_t *mask,
                                 LineGeom g, const float *k, int R, bool bf, hipStream_t s) {
    plan_segments_lowp(g);
    g.vec_out = (g.out_ax == 1 && g.out_lane % 4 == 0 && g.out_outer % 4 == 0 &&
                 reinterpret_cast<uintptr_t>(out) % 8 == 0 &&
                 (mask == nullptr || reinterpret_cast<uintptr_t>(mask) % 4 == 0)) ? 1 : 0;
    const int64_t blocks = g.nseg * ((g.n_lane + 63) / 64) * g.n_outer;
    if (blocks == 0) return hipSuccess;
    const TapsL<2 * RT + 1> t = split_taps<RT>(k, R, bf);
    if (bf)
        hipLaunchKernelGGL((k_line_lowp<RT, true>), dim3((unsigned)blocks), dim3(256), 0, s, in, out, mask, g, t);
    else
        hipLaunchKernelGGL((k_line_lowp<RT, false>), dim3((unsigned)blocks), dim3(256), 0, s, in, out, mask, g, t);
    return hipGetLastError();
}

static bool lowp_offsets_fit(const LineGeom &g) {
    return (g.in_ax * 2 * 31 + g.n_lane * 2) < (int64_t(1) << 32);
}

hipError_t launch_line_pass_lowp(const uint16_t *in, uint16_t *out, const uint8_t *mask,
                                 const LineGeom &g, const float *taps, int R, bool bf,
                                 hipStream_t s, MaskBits bits) {
    if (!force_ring_kernels() && lowp_lds_supported(in, g, R))
        return launch_line_pass_lowp_lds(in, out, mask, bits, g, taps, R, bf, s);
    if (!lowp_offsets_fit(g)) return hipErrorInvalidValue;
#define MXB_LOWP_CASE(rt) \
    if (R <= rt) return launch_lowp_rt<rt>(in, out, mask, g, taps, R, bf, s);
    MXB_LOWP_CASE(4)
    MXB_LOWP_CASE(8)
    MXB_LOWP_CASE(12)
    MXB_LOWP_CASE(16)
    MXB_LOWP_CASE(24)
    MXB_LOWP_CASE(32)
    MXB_LOWP_CASE(40)
    MXB_LOWP_CASE(48)
    MXB_LOWP_CASE(64)
#undef MXB_LOWP_CASE
    return hipErrorInvalidValue;
}

static inline int64_t blocks_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    return b < 1 ? 1 : (b > 65536 ? 65536 : b);
}

// Whether every pointer is aligned to its vector width (mask 8, the rest 16).
static inline bool vec_ok(const void *mask8, const void *a16, const void *b16 = nullptr) {
    return reinterpret_cast<uintptr_t>(mask8) % 8 == 0 && reinterpret_cast<uintptr_t>(a16) % 16 == 0 &&
           reinterpret_cast<uintptr_t>(b16) % 16 == 0;
}

hipError_t launch_init_lowp(const uint8_t *mask, uint16_t *u, int64_t n, bool bf, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const int vec = vec_ok(mask, u);
    const dim3 grid((unsigned)blocks_for(vec ? (n + 7) / 8 : n));
    if (bf)
        hipLaunchKernelGGL(k_init_lowp<true>, grid, dim3(256), 0, s, mask, u, n, vec);
    else
        hipLaunchKernelGGL(k_init_lowp<false>, grid, dim3(256), 0, s, mask, u, n, vec);
    return hipGetLastError();
}

hipError_t launch_lowp_to_f32(const uint16_t *u, float *f, int64_t n, bool bf, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const int vec = vec_ok(nullptr, u, f);
    const dim3 grid((unsigned)blocks_for(vec ? (n + 7) / 8 : n));
    if (bf)
        hipLaunchKernelGGL(k_lowp_to_f32<true>, grid, dim3(256), 0, s, u, f, n, vec);
    else
        hipLaunchKernelGGL(k_lowp_to_f32<false>, grid, dim3(256), 0, s, u, f, n, vec);
    return hipGetLastError();
}

}  // namespace mxb

namespace mxb {

hipError_t launch_init_split(const uint8_t *mask, uint16_t *planes, int64_t lo_off, int64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const int vec = vec_ok(mask, planes, planes + lo_off);
    hipLaunchKernelGGL(k_init_split, dim3((unsigned)blocks_for(vec ? (n + 7) / 8 : n)), dim3(256), 0, s, mask,
                       planes, lo_off, n, vec);
    return hipGetLastError();
}

hipError_t launch_split_state(const float *u, uint16_t *planes, int64_t lo_off, int64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_split_state, dim3((unsigned)blocks_for(n)), dim3(256), 0, s, u, planes, lo_off, n);
    return hipGetLastError();
}

}  // namespace mxb
