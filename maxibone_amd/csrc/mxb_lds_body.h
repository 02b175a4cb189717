// FP32 line pass, LDS-staged form (gfx950): the per-block body, shared by the
// single-volume kernel (mxb_kernels_f32_lds.hip) and the batched persistent
// kernel (mxb_batched.hip).
//
// Same arithmetic as k_line_mfma (banded-Toeplitz on v_mfma_f32_16x16x4_f32,
// FP32 accumulation, zero outside the line) but the inputs travel
// global -> VGPR -> LDS in 16-byte-per-lane loads: one block-wide load
// instruction moves 16 positions x 64 lines = 16 full 256-byte rows, instead of
// the 64-byte pieces the per-lane MFMA operand layout would fetch directly.
//
// Block = 256 threads = 4 waves over 64 consecutive lines; wave w computes
// lines 16w..16w+15.  Positions are grouped in slots of 16 (one tile): slot s
// holds positions a0 - RT + 16 s .. +15, stored as rows of 64 floats.  A tile
// t (outputs a0 + 16t .. +15) reads W = ceil((16 + 2RT)/16) slots, t .. t+W-1.
// The LDS ring holds NSL = W + 1 slots: while tile t computes, slot t + W is
// being written from the VGPRs loaded one tile earlier, and the global loads
// for slot t + W + 1 are in flight.
//
// Rows are stored rotated by 16 floats per position (row p holds line l at
// column (l + 16 (p mod 4)) mod 64), so the 4 rows x 16 lines an MFMA operand
// read touches fall in distinct banks.
#pragma once
#include "mxb_internal.h"

namespace mxb {

typedef float v4f __attribute__((ext_vector_type(4)));

template <int N>
struct TapsF {
    float k[N];
};


template <int RT>
struct LdsShape {
    static constexpr int C = (16 + 2 * RT + 3) / 4;   // MFMA chunks per tile
    static constexpr int W = (16 + 2 * RT + 15) / 16; // slots a tile reads
    static constexpr int NSL = W + 1;                 // LDS ring slots
};

__device__ __forceinline__ uint16_t quantise_u16_l(float v) {
#pragma clang fp contract(off)
    const float s = v * 65535.0f;
    return (uint16_t)(int32_t)floorf(s);
}

// Shared-memory layout of one block: [ring | tap table | stage].
template <int RT>
struct LdsSmem {
    static constexpr int RING = LdsShape<RT>::NSL * 16 * 64;
    static constexpr int TAPS = 4 * LdsShape<RT>::C + 32;
    static constexpr int STAGE = 64 * 36;  // 16 x 64 (lines contiguous) or 64 x 36 (positions)
    static constexpr int TOTAL = RING + TAPS + STAGE;
};

// Fill the zero-padded tap table of `smem`: entry t + 16 holds k[t],
// t in [0, 2RT], from the centred taps `k` (2RT + 1 values, any address space).
template <int RT, typename TapSrc>
__device__ __forceinline__ void lds_fill_taps(float *smem, const TapSrc &k) {
    float *tap_lds = smem + LdsSmem<RT>::RING;
    for (int i = threadIdx.x; i < LdsSmem<RT>::TAPS; i += 256) {
        const int t = i - 16;
        tap_lds[i] = (t >= 0 && t <= 2 * RT) ? k[t] : 0.0f;
    }
}

// One block's share of one line pass: 64 lines starting at group `lb` of
// outer index `outer`, positions of segment `seg_id`.  The tap table must have
// been filled (lds_fill_taps) by the calling block; the body synchronises
// before reading it.  Returns with every thread past a final barrier, so the
// caller may reuse `smem`.
template <int RT>
__device__ __forceinline__ void lds_line_body(const float *__restrict__ in, const LineGeom &g,
                                              const Epilogue &e, int64_t outer, int64_t lb,
                                              int64_t seg_id, float *smem) {
    using S = LdsShape<RT>;
    constexpr int C = S::C, W = S::W, NSL = S::NSL;
    float *ring = smem;
    float *tap_lds = smem + LdsSmem<RT>::RING;
    float *stage = smem + LdsSmem<RT>::RING + LdsSmem<RT>::TAPS;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int col = lane & 15, kg = lane >> 4;

    const int64_t line0 = lb * 64;
    const int64_t a0 = seg_id * g.seg;
    const int64_t a1 = a0 + g.seg < g.n_ax ? a0 + g.seg : g.n_ax;
    const int64_t n_ax = g.n_ax;

    // ---- loader mapping: thread -> (row r = tid / 16, 4 lines at 4 * (tid % 16))
    const int lr = tid >> 4;
    const int lq = tid & 15;
    const int64_t lline = line0 + 4 * lq;
    // lines beyond n_lane read line n_lane-1 (their outputs are never stored);
    // a quad that straddles n_lane is loaded element-wise
    const bool quad_full = lline + 3 < g.n_lane;
    const char *lbase = reinterpret_cast<const char *>(in + outer * g.in_outer);
    const int64_t in_ax_b = g.in_ax * 4;
    const int64_t pos_first = a0 - RT;  // position of slot 0, row 0

    auto gload = [&](int64_t slot) -> v4f {
        const int64_t p = pos_first + 16 * slot + lr;
        if (p < 0 || p >= n_ax) return v4f{0.0f, 0.0f, 0.0f, 0.0f};
        const char *rowp = lbase + p * in_ax_b;
        if (quad_full) return *reinterpret_cast<const v4f *>(rowp + lline * 4);
        v4f v;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t l = lline + q < g.n_lane ? lline + q : g.n_lane - 1;
            v[q] = *reinterpret_cast<const float *>(rowp + l * 4);
        }
        return v;
    };
    // LDS word offset of (slot ring index si, row r, line l): rotated rows
    auto lds_off = [](int si, int r, int l) -> int {
        return si * 1024 + r * 64 + ((l + 16 * (r & 3)) & 63);
    };
    auto lwrite = [&](int64_t slot, v4f v) {
        const int si = (int)(slot % NSL);
        *reinterpret_cast<v4f *>(&ring[lds_off(si, lr, 4 * lq)]) = v;
    };

    // Fill the first W slots, stage slot W in VGPRs.
    for (int s = 0; s < W; ++s) lwrite(s, gload(s));
    v4f staged = gload(W);
    __syncthreads();

    float tv[C];
#pragma unroll
    for (int c = 0; c < C; ++c) tv[c] = tap_lds[4 * c + kg - col + 16];

    // per-lane LDS read offset within a slot for row group kg, line 16w + col
    const int rd_lane = kg * 64 + ((16 * wv + col + 16 * kg) & 63);
    const int64_t li = line0 + 16 * wv + col;
    const bool act = li < g.n_lane;
    const bool zpass_stage = g.stage_out != 0;
    const int64_t ntiles = (a1 - a0 + 15) / 16;

    for (int64_t t = 0; t < ntiles; ++t) {
        // 1. slot t + W (staged last iteration) into the ring slot freed by tile t-1
        lwrite(t + W, staged);
        // 2. global loads for slot t + W + 1 (consumed next iteration)
        if (t + 1 < ntiles) staged = gload(t + W + 1);
        // 3. compute tile t from slots t .. t + W - 1
        int sb[W];
#pragma unroll
        for (int s = 0; s < W; ++s) sb[s] = (int)((t + s) % NSL) * 1024;
        float x[C];
#pragma unroll
        for (int c = 0; c < C; ++c) x[c] = ring[sb[c / 4] + (c % 4) * 4 * 64 + rd_lane];
        // One accumulation chain in tap order: the result then does not depend
        // on where the tile grid falls along the line (chunked and in-memory
        // runs agree bit for bit).
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c = 0; c < C; ++c) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(tv[c], x[c], acc, 0, 0, 0);
        // 4. epilogue: outputs a0 + 16t + 4kg + r of line li
        const int64_t tpos = a0 + 16 * t;
        if (zpass_stage) {
            // lines contiguous in the output: stage the 16 x 64 tile, then
            // every thread stores 4 lines of one position (256-byte rows)
#pragma unroll
            for (int r = 0; r < 4; ++r) stage[(4 * kg + r) * 64 + 16 * wv + col] = acc[r];
            __syncthreads();
            const int64_t p = tpos + lr;
            if (p < a1) {
                const v4f v = *reinterpret_cast<const v4f *>(&stage[lr * 64 + 4 * lq]);
                float *op = e.out_f32 + outer * g.out_outer + p * g.out_ax;
                if (quad_full) {
                    *reinterpret_cast<v4f *>(op + lline) = v;
                } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (lline + q < g.n_lane) op[lline + q] = v[q];
                }
            }
        } else if (g.vec_out) {
            // positions contiguous in the output: collect two tiles (32
            // positions) per line, then write whole 128-byte line pieces
            const int half = (int)(t & 1);
            *reinterpret_cast<v4f *>(&stage[(16 * wv + col) * 36 + 16 * half + 4 * kg]) = acc;
            if (half == 1 || t == ntiles - 1) {
                __syncthreads();
                const int64_t tp = tpos - 16 * half;  // first position of the pair
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const int line = 32 * k + (tid >> 3);
                    const int q = tid & 7;
                    const int64_t gl = line0 + line;
                    const int64_t p0 = tp + 4 * q;
                    if (gl < g.n_lane && p0 < a1) {
                        v4f v = *reinterpret_cast<const v4f *>(&stage[line * 36 + 4 * q]);
                        const int64_t off = outer * g.out_outer + gl * g.out_lane + p0;
                        if (p0 + 3 < a1) {
                            if (e.mask != nullptr) {
                                const uint32_t m = *reinterpret_cast<const uint32_t *>(e.mask + off);
#pragma unroll
                                for (int r = 0; r < 4; ++r)
                                    if ((m >> (8 * r)) & 0xffu) v[r] = 1.0f;
                            }
                            if (e.out_f32 != nullptr) *reinterpret_cast<v4f *>(e.out_f32 + off) = v;
                            if (e.out_u16 != nullptr) {
                                const uint32_t lo = (uint32_t)quantise_u16_l(v[0]) | ((uint32_t)quantise_u16_l(v[1]) << 16);
                                const uint32_t hi = (uint32_t)quantise_u16_l(v[2]) | ((uint32_t)quantise_u16_l(v[3]) << 16);
                                *reinterpret_cast<uint2 *>(e.out_u16 + off) = make_uint2(lo, hi);
                            }
                        } else {
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                if (p0 + r >= a1) break;
                                float w = v[r];
                                if (e.mask != nullptr && e.mask[off + r] != 0) w = 1.0f;
                                if (e.out_f32 != nullptr) e.out_f32[off + r] = w;
                                if (e.out_u16 != nullptr) e.out_u16[off + r] = quantise_u16_l(w);
                            }
                        }
                    }
                }
            }
        } else if (act) {
            const int64_t obase = outer * g.out_outer + li * g.out_lane;
            const int64_t p0 = tpos + 4 * kg;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (p0 + r >= a1) break;
                const int64_t off = obase + (p0 + r) * g.out_ax;
                float v = acc[r];
                if (e.mask != nullptr && e.mask[off] != 0) v = 1.0f;
                if (e.out_f32 != nullptr) e.out_f32[off] = v;
                if (e.out_u16 != nullptr) e.out_u16[off] = quantise_u16_l(v);
            }
        }
        // all waves are done with slot t (and with `stage`) before it is reused
        __syncthreads();
    }
}


}  // namespace mxb
