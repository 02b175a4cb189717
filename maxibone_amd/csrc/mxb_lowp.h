// Pieces shared by the two reduced-precision (BF16 / FP16 state) pass kernels:
// the 16-bit format traits, the hi/lo tap split and the tile shape.
//
// Each pass is a banded-Toeplitz contraction on
// v_mfma_f32_16x16x32_{bf16,f16} with FP32 accumulation:
//   D[i][j] += T_c[i][kk] * X_c[kk][j],  i = output (16), j = line (16),
//   kk = input position within a 32-position chunk c,
//   T_c[i][kk] = k[32c + kk - i] (zero outside the band),
// where tile t (outputs a0 + 16t .. +15) takes its chunk c from input
// positions a0 - RT + 16t + 32c .. +31.  Each tap is split into hi + lo
// 16-bit parts (two MFMAs per chunk, hi first), so the taps carry ~16
// significant bits and their sum stays 1 to ~2^-17: the 16-bit rounding
// happens only where the state is stored.
//
// Lane layout of the 16x16x32 operands (lane l, g = l / 16, j = l % 16):
//   A (taps):   T_c[i = l%16][kk = 8g .. 8g+7]
//   B (inputs): X_c[kk = 8g .. 8g+7][j = l%16]
//   D:          out_j(tile base + 4g + r), r = 0..3
#pragma once
#include <cmath>
#include <cstring>

#include "mxb_internal.h"

namespace mxb {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

template <bool BF>
struct Lowp;

template <>
struct Lowp<true> {
    static __device__ __forceinline__ v4f mfma(v8s a, v8s b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a),
                                                       __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
    }
    static __device__ __forceinline__ uint16_t from_f32(float v) {
        return __builtin_bit_cast(uint16_t, (__bf16)v);  // round to nearest even
    }
    static __device__ __forceinline__ float to_f32(uint16_t v) {
        return __builtin_bit_cast(float, (uint32_t)v << 16);
    }
};

template <>
struct Lowp<false> {
    static __device__ __forceinline__ v4f mfma(v8s a, v8s b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(v8h, a),
                                                      __builtin_bit_cast(v8h, b), c, 0, 0, 0);
    }
    static __device__ __forceinline__ uint16_t from_f32(float v) {
        return __builtin_bit_cast(uint16_t, (_Float16)v);
    }
    static __device__ __forceinline__ float to_f32(uint16_t v) {
        return (float)__builtin_bit_cast(_Float16, v);
    }
};

// floor(v * 65535.0f) as uint16, through int32 (cpu:191-196).
__device__ __forceinline__ uint16_t quantise_u16_s(float v) {
#pragma clang fp contract(off)
    const float s = v * 65535.0f;
    return (uint16_t)(int32_t)floorf(s);
}

// FP16 high and low parts of v for the split (FP32-class) form:
// hi = f16(v), lo = f16(v - hi), where v - hi is exact in FP32.
__device__ __forceinline__ void split_f16(float v, uint16_t &hi, uint16_t &lo) {
#pragma clang fp contract(off)
    const _Float16 h = (_Float16)v;
    hi = __builtin_bit_cast(uint16_t, h);
    lo = __builtin_bit_cast(uint16_t, (_Float16)(v - (float)h));
}

// Eight BF16 values (one MFMA operand) as FP16: exact from 2^-14, FP16's
// smallest normal number, up to 65504; smaller magnitudes round to FP16's
// subnormal step of 2^-24 (an error below 2^-25).
__device__ __forceinline__ v8s bf16x8_to_f16(v8s x) {
    v8s r;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float f = __builtin_bit_cast(float, (uint32_t)(uint16_t)x[e] << 16);
        r[e] = __builtin_bit_cast(short, (_Float16)f);
    }
    return r;
}

// split_f16(a * scale) for a power-of-two `scale`, where the product is
// exact: the low part's difference a * scale - hi comes from one FMA, which
// rounds the same exact value once, so the parts equal split_f16's (one
// instruction fewer than multiplying first).
__device__ __forceinline__ void split_f16_scaled(float a, float scale, uint16_t &hi, uint16_t &lo) {
#pragma clang fp contract(off)
    const _Float16 h = (_Float16)(a * scale);
    hi = __builtin_bit_cast(uint16_t, h);
    lo = __builtin_bit_cast(uint16_t, (_Float16)__builtin_fmaf(a, scale, -(float)h));
}

__device__ __forceinline__ uint32_t pack2(uint16_t lo, uint16_t hi) {
    return (uint32_t)lo | ((uint32_t)hi << 16);
}

// 16-bit tap bit patterns, hi and lo parts, centred in a capacity of 2RT+1.
template <int N>
struct TapsL {
    uint16_t hi[N];
    uint16_t lo[N];
};

template <int RT>
struct LowpChunks {
    static constexpr int C = (16 + 2 * RT + 31) / 32;  // 32-position chunks per tile
    static constexpr int TW = 32 * C + 32;             // zero-padded tap table width
};

// Fill the zero-padded tap tables, t*[x + 16] = k[x] for x in [0, 2RT], and
// (after the caller's barrier) read a lane's A operands from them:
// T_c[col][8kg + e] = k[32c + 8kg + e - col].
template <int RT>
__device__ __forceinline__ void lowp_fill_taps(uint16_t *thi, uint16_t *tlo,
                                               const TapsL<2 * RT + 1> &taps) {
    for (int i = threadIdx.x; i < LowpChunks<RT>::TW; i += blockDim.x) {
        const int t = i - 16;
        const bool in_band = t >= 0 && t <= 2 * RT;
        thi[i] = in_band ? taps.hi[in_band ? t : 0] : 0;
        tlo[i] = in_band ? taps.lo[in_band ? t : 0] : 0;
    }
}

template <int RT>
__device__ __forceinline__ void lowp_read_taps(const uint16_t *thi, const uint16_t *tlo, int col,
                                               int kg, v8s *ahi, v8s *alo) {
#pragma unroll
    for (int c = 0; c < LowpChunks<RT>::C; ++c)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int x = 32 * c + 8 * kg + e - col + 16;
            ahi[c][e] = (short)thi[x];
            alo[c][e] = (short)tlo[x];
        }
}

// ---------------------------------------------------------------- host side
inline uint16_t bf16_bits(float v) {  // round to nearest even
    uint32_t x;
    std::memcpy(&x, &v, 4);
    const uint32_t lsb = (x >> 16) & 1u;
    x += 0x7fffu + lsb;
    return (uint16_t)(x >> 16);
}

inline float bf16_value(uint16_t b) {
    const uint32_t x = (uint32_t)b << 16;
    float v;
    std::memcpy(&v, &x, 4);
    return v;
}

inline uint16_t f16_bits(float v) {
    const _Float16 h = (_Float16)v;
    uint16_t b;
    std::memcpy(&b, &h, 2);
    return b;
}

inline float f16_value(uint16_t b) {
    _Float16 h;
    std::memcpy(&h, &b, 2);
    return (float)h;
}

// Taps of radius R <= RT, centred in 2RT+1 entries, split into hi + lo parts.
template <int RT>
TapsL<2 * RT + 1> split_taps(const float *k, int R, bool bf) {
    TapsL<2 * RT + 1> t;
    for (int i = 0; i < 2 * RT + 1; ++i) {
        const int r = i - RT;
        const float v = (r >= -R && r <= R) ? k[R + r] : 0.0f;
        const uint16_t hi = bf ? bf16_bits(v) : f16_bits(v);
        const float rest = v - (bf ? bf16_value(hi) : f16_value(hi));
        t.hi[i] = hi;
        t.lo[i] = bf ? bf16_bits(rest) : f16_bits(rest);
    }
    return t;
}

// Taps of the split form: k * 2^shift (shift chosen so that the largest |tap|
// lies in [2^14, 2^15) and the low parts stay normal FP16 numbers), split into
// FP16 hi + lo; *scale = 2^-shift undoes the scaling of an accumulator exactly.
template <int RT>
TapsL<2 * RT + 1> split_taps_scaled(const float *k, int R, float *scale) {
    float amax = 0.0f;
    for (int i = 0; i < 2 * R + 1; ++i) amax = std::fmax(amax, std::fabs(k[i]));
    int e = 0;
    if (amax > 0.0f) (void)std::frexp(amax, &e);  // amax in [2^(e-1), 2^e)
    const int shift = amax > 0.0f ? 15 - e : 0;
    float kk[2 * 64 + 1];
    for (int i = 0; i < 2 * R + 1; ++i) kk[i] = std::ldexp(k[i], shift);
    *scale = std::ldexp(1.0f, -shift);
    return split_taps<RT>(kk, R, false);
}

// LDS-staged form (mxb_kernels_lowp_lds.hip): usable when every input row is
// 16-byte aligned and the radius is compiled.
bool lowp_lds_supported(const uint16_t *in, const LineGeom &g, int R);
hipError_t launch_line_pass_lowp_lds(const uint16_t *in, uint16_t *out, const uint8_t *mask,
                                     const MaskBits &bits, const LineGeom &g, const float *taps,
                                     int R, bool bf, hipStream_t s);

}  // namespace mxb
