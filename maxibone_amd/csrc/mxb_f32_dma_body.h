// FP32 line pass, direct-load form (gfx950): the per-block bodies, shared by
// the single-volume kernel (mxb_kernels_f32_lds.hip) and the batched
// persistent kernel (mxb_batched.hip).  f32_dma_body runs one 16-position
// tile per iteration (the line-contiguous z-pass), f32_dma_body2 two (the
// position-contiguous y and x passes, the x-pass illuminating from mask
// bits); f32_dma_pass picks per epilogue form.
#pragma once
#include "mxb_dma.h"
#include "mxb_lds_body.h"

namespace mxb {

// Direct-load form: the same pass and the same LDS row layout, with each slot
// (16 positions x 64 lines, 4 KB) moved global -> LDS by
// global_load_lds_dwordx4, D slots ahead of use, with no VGPR staging.
// Lane i of wave w writes LDS bytes 1024 w + 16 i of its slot: row
// r = 4w + i/16 at float column 4 (i % 16), which holds lines
// (4 (i % 16) - 16 (r mod 4)) mod 64 .. +3 of the rotated row.
//
// Iteration j (tile j): A. wait until group j has landed, barrier;
// B. issue group j + D into the slots freed by tile j - 1; C. compute tile j;
// D. stage the outputs in LDS, barrier, S stores per thread.  Group i = data
// slot i + W - 1; the wait count is (D-1) + D S once j >= D and D-1 before,
// as in the 16-bit direct-load pass.
template <int RT, int D>
struct F32DmaShape {
    static constexpr int C = LdsShape<RT>::C;
    static constexpr int W = LdsShape<RT>::W;
    static constexpr int NSL = W + D;  // ring slots
};

#ifndef MXB_F32_DMA_DEPTH
#define MXB_F32_DMA_DEPTH 2
#endif
// Pair depth of the two-tile body: the y-pass and the illuminating x-pass
// (whose D - 1 must be a multiple of 4).
#ifndef MXB_F32_DEPTH_Y
#define MXB_F32_DEPTH_Y 1
#endif
#ifndef MXB_F32_DEPTH_X
#define MXB_F32_DEPTH_X 1
#endif

constexpr int kF32Slot = 16 * 64;        // floats per slot
constexpr int kF32StagePitchV = 36;      // floats per line row (position-contiguous epilogue, 2 tiles)
// floats per position row of the line-contiguous epilogue: 68 puts the four
// row groups one MFMA result writes (4 kg + r) in distinct 16-bank groups
constexpr int kF32StagePitchZ = 68;

// Shared-memory layout of one block of the one-tile pass (bytes):
// [ring | stage | tap table].
template <int RT, int D, bool MASK, bool ZSTAGE>
struct F32DmaSmem {
    static_assert(!MASK, "the illuminating pass runs on f32_dma_body2 (mask bits)");
    using S = F32DmaShape<RT, D>;
    static constexpr int RING = S::NSL * kF32Slot * 4;
    static constexpr int STAGE = (ZSTAGE ? 16 * kF32StagePitchZ : 64 * kF32StagePitchV) * 4;
    static constexpr int TAPS = (4 * S::C + 32) * 4;
    static constexpr int BYTES = RING + STAGE + TAPS;
};

// One block's share of one pass: 64 lines from group `lb` of outer index
// `outer`, positions of segment `seg_id`; `taps` holds the 2RT+1 centred taps
// (any address space).  Returns with the wave's loads and stores drained and
// every thread past a final barrier, so the caller may reuse `smem` at once.
template <int RT, int D, bool MASK, bool OUTF, bool OUTU, bool ZSTAGE, typename TapSrc>
__device__ __forceinline__ void f32_dma_body(const float *__restrict__ in, const LineGeom &g,
                                             const Epilogue &e, int64_t outer, int64_t lb,
                                             int64_t seg_id, const TapSrc &taps,
                                             unsigned char *__restrict__ smem) {
    using S = F32DmaShape<RT, D>;
    constexpr int C = S::C, W = S::W, NSL = S::NSL;
    // stores per thread: every tile (line-contiguous) or every second tile
    // (position-contiguous: a pair of tiles makes 128-byte line pieces)
    constexpr int ST = ZSTAGE ? 1 : (OUTF ? 2 : 0) + (OUTU ? 1 : 0);
    static_assert(ZSTAGE || D % 2 == 0, "pair flushes keep the store count fixed only for even D");
    static_assert(!MASK, "the illuminating pass runs on f32_dma_body2 (mask bits)");
    static_assert(!ZSTAGE || (OUTF && !OUTU && !MASK), "line-contiguous epilogue writes floats only");
    static_assert(ST >= 1, "at least one output");
    using M = F32DmaSmem<RT, D, MASK, ZSTAGE>;
    constexpr int TAPS = M::TAPS / 4;
    float *const ring = reinterpret_cast<float *>(smem);
    float *const stage = reinterpret_cast<float *>(smem + M::RING);
    float *const tap_lds = reinterpret_cast<float *>(smem + M::RING + M::STAGE);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int col = lane & 15, kg = lane >> 4;
    const int64_t line0 = lb * 64;
    const int64_t a0 = seg_id * g.seg;
    const int64_t a1 = a0 + g.seg < g.n_ax ? a0 + g.seg : g.n_ax;
    const int64_t n_ax = g.n_ax;
    const int ntiles = (int)((a1 - a0 + 15) / 16);
    const int nslots = ntiles + W - 1;

    for (int i = tid; i < TAPS; i += 256) {
        const int t = i - 16;
        tap_lds[i] = (t >= 0 && t <= 2 * RT) ? taps[t] : 0.0f;
    }

    // ---- data loads
    const int dr = 4 * wv + (lane >> 4);
    const int dl = (4 * (lane & 15) - 16 * (dr & 3)) & 63;
    const bool dl_ok = line0 + dl < g.n_lane;  // n_lane is a multiple of 4
    const char *dsrc = reinterpret_cast<const char *>(in + outer * g.in_outer + line0 + dl);
    const int64_t in_ax_b = g.in_ax * 4;
    const int64_t pos_first = a0 - RT;
    const uint32_t ring_lds = lds_addr(ring) + 1024u * (uint32_t)wv;
    auto issue_data = [&](int slot) {
        const int64_t p = pos_first + 16 * (int64_t)slot + dr;
        const bool ok = dl_ok && slot < nslots && p >= 0 && p < n_ax;
        const void *src = ok ? static_cast<const void *>(dsrc + p * in_ax_b)
                             : static_cast<const void *>(g_dma_zero);
        dma16<ZSTAGE>(ring_lds + (uint32_t)((slot % NSL) * kF32Slot * 4), src);
    };
    auto issue_group = [&](int i) { issue_data(i + W - 1); };

    for (int s = 0; s < W - 1; ++s) issue_data(s);
    for (int i = 0; i < D; ++i) issue_group(i);
    __syncthreads();  // tap table

    float tv[C];
#pragma unroll
    for (int c = 0; c < C; ++c) tv[c] = tap_lds[4 * c + kg - col + 16];
    const int rd_lane = kg * 64 + ((16 * wv + col + 16 * kg) & 63);

    // store-side constants
    const int zr = tid >> 4, zq = tid & 15;  // line-contiguous: position, 4 lines
    const bool zq_ok = line0 + 4 * zq < g.n_lane;
    const int sl = tid >> 2, sq = tid & 3;   // position-contiguous: line, 8 positions
    const bool sl_ok = line0 + sl < g.n_lane;
    float *const sinkf = reinterpret_cast<float *>(g_dma_sink + lane);
    uint16_t *const sinku = reinterpret_cast<uint16_t *>(g_dma_sink + lane);

    for (int j = 0; j < ntiles; ++j) {
        if (j >= D)
            vm_wait<(D - 1) + (ZSTAGE ? D : D / 2) * ST>();
        else
            vm_wait<D - 1>();
        __syncthreads();
        issue_group(j + D);
        int sb[W];
#pragma unroll
        for (int s2 = 0; s2 < W; ++s2) sb[s2] = ((j + s2) % NSL) * kF32Slot;
        float x[C];
#pragma unroll
        for (int c = 0; c < C; ++c) x[c] = ring[sb[c / 4] + (c % 4) * 256 + rd_lane];
        // one accumulation chain in tap order (bit-identical to the
        // one-FMA-per-tap form, whatever the tile alignment)
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c = 0; c < C; ++c) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(tv[c], x[c], acc, 0, 0, 0);
        const int64_t tpos = a0 + 16 * (int64_t)j;
        const bool full = tpos + 16 <= a1;  // uniform; only a segment's last tile is partial
        if (ZSTAGE) {
#pragma unroll
            for (int r = 0; r < 4; ++r) stage[(4 * kg + r) * kF32StagePitchZ + 16 * wv + col] = acc[r];
            __syncthreads();
            const v4f v = *reinterpret_cast<const v4f *>(&stage[zr * kF32StagePitchZ + 4 * zq]);
            const int64_t p = tpos + zr;
            float *op = e.out_f32 + outer * g.out_outer + p * g.out_ax + line0 + 4 * zq;
            if (full) {
                if (nt_rows_ok(e.out_f32, 4 * g.out_outer))
                    store16<true>(zq_ok ? op : sinkf, v);
                else
                    store16<false>(zq_ok ? op : sinkf, v);
            } else if (zq_ok && p < a1) {
                store16<false>(op, v);
            }
        } else {
            const int half = j & 1;
            *reinterpret_cast<v4f *>(&stage[(16 * wv + col) * kF32StagePitchV + 16 * half + 4 * kg]) = acc;
            if (half == 1 || j == ntiles - 1) {
                __syncthreads();
                const int64_t ppos = tpos - 16 * half;  // first position of the pair
                const bool pfull = ppos + 32 <= a1;     // uniform
                v4f v[2];
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    v[h] = *reinterpret_cast<const v4f *>(&stage[sl * kF32StagePitchV + 8 * sq + 4 * h]);
                const int64_t p0 = ppos + 8 * sq;
                const int64_t off = outer * g.out_outer + (line0 + sl) * g.out_lane + p0;
                if (pfull && half == 1) {
                    if (OUTF) {
                        float *dst = sl_ok ? e.out_f32 + off : sinkf;
                        store16<false>(dst, v[0]);
                        store16<false>(dst + (sl_ok ? 4 : 0), v[1]);
                    }
                    if (OUTU) {
                        uint32_t w[4];
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            w[2 * h] = (uint32_t)quantise_u16_l(v[h][0]) | ((uint32_t)quantise_u16_l(v[h][1]) << 16);
                            w[2 * h + 1] = (uint32_t)quantise_u16_l(v[h][2]) | ((uint32_t)quantise_u16_l(v[h][3]) << 16);
                        }
                        store16<false>(sl_ok ? e.out_u16 + off : sinku, make_uint4(w[0], w[1], w[2], w[3]));
                    }
                } else if (sl_ok) {
                    // a segment's last (partial or unpaired) tiles
#pragma unroll
                    for (int r = 0; r < 8; ++r) {
                        if (p0 + r >= a1 || p0 + r >= tpos + 16) break;
                        const float w = v[r >> 2][r & 3];
                        if (OUTF) gmem(e.out_f32)[off + r] = w;
                        if (OUTU) gmem(e.out_u16)[off + r] = quantise_u16_l(w);
                    }
                }
            }
        }
    }
    // drain this wave's loads (prefetches past the segment land in the ring)
    // and stores, and let every wave finish with the LDS
    vm_wait<0>();
    __syncthreads();
}


// ---------------------------------------------------------------------------
// Two tiles per iteration (positions 32j .. 32j+31 of the segment): the same
// slots, layout and arithmetic, with the barrier, epilogue and load issue of
// an iteration amortised over twice the MFMA work.  Tile 1's operand rows are
// tile 0's shifted by 16 positions = 4 chunks, so one set of C + 4 LDS reads
// feeds both accumulation chains (each chain still adds its taps in order:
// the results are bit-identical to the one-tile form).
//
// Group i = data slots 2i + W - 1, 2i + W (+, when i % 4 == 0, the mask words
// of pairs i .. i+3).  Pair j reads slots 2j .. 2j + W; the ring holds
// W + 1 + 2D slots; the wait before pair j is vmcnt(GW + D S) once j >= D,
// with GW the loads of D - 1 groups and S stores per pair.
template <int RT, int D>
struct F32Dma2Shape {
    static constexpr int C = LdsShape<RT>::C;
    static constexpr int W = LdsShape<RT>::W;
    static constexpr int NSL = W + 1 + 2 * D;
};

template <int RT, int D, bool MASK, bool ZSTAGE>
struct F32Dma2Smem {
    using S = F32Dma2Shape<RT, D>;
    static constexpr int RING = S::NSL * kF32Slot * 4;
    static constexpr int NM4 = (D + 7) / 4;  // mask slots of 4 pairs: reused only after pair i+3's epilogue
    static constexpr int MRING = MASK ? NM4 * 1024 : 0;  // 64 lines x 4 pairs' 32-bit mask words
    static constexpr int STAGE = (ZSTAGE ? 32 * kF32StagePitchZ : 64 * kF32StagePitchV) * 4;
    static constexpr int TAPS = (4 * S::C + 32) * 4;
    static constexpr int BYTES = RING + MRING + STAGE + TAPS;
};

template <int RT, int D, bool MASK, bool OUTF, bool OUTU, bool ZSTAGE, typename TapSrc>
__device__ __forceinline__ void f32_dma_body2(const float *__restrict__ in, const LineGeom &g,
                                              const Epilogue &e, int64_t outer, int64_t lb,
                                              int64_t seg_id, const TapSrc &taps,
                                              unsigned char *__restrict__ smem) {
    using S = F32Dma2Shape<RT, D>;
    constexpr int C = S::C, W = S::W, NSL = S::NSL;
    // loads issued in D - 1 groups: 2 data slots each, and (with (D - 1) % 4
    // == 0) exactly (D - 1) / 4 mask loads (one per 4 pairs)
    constexpr int GW = 2 * (D - 1) + (MASK ? (D - 1) / 4 : 0);
    static_assert(!MASK || (D - 1) % 4 == 0, "a fixed count of mask loads per D - 1 groups");
    constexpr int NM4 = F32Dma2Smem<RT, D, MASK, ZSTAGE>::NM4;
    constexpr int ST = ZSTAGE ? 2 : (OUTF ? 2 : 0) + (OUTU ? 1 : 0);
    static_assert(!ZSTAGE || (OUTF && !OUTU && !MASK), "line-contiguous epilogue writes floats only");
    static_assert(ST >= 1, "at least one output");
    using M = F32Dma2Smem<RT, D, MASK, ZSTAGE>;
    constexpr int TAPS = M::TAPS / 4;
    float *const ring = reinterpret_cast<float *>(smem);
    unsigned char *const mring = smem + M::RING;
    float *const stage = reinterpret_cast<float *>(smem + M::RING + M::MRING);
    float *const tap_lds = reinterpret_cast<float *>(smem + M::RING + M::MRING + M::STAGE);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int col = lane & 15, kg = lane >> 4;
    const int64_t line0 = lb * 64;
    const int64_t a0 = seg_id * g.seg;
    const int64_t a1 = a0 + g.seg < g.n_ax ? a0 + g.seg : g.n_ax;
    const int64_t n_ax = g.n_ax;
    const int npairs = (int)((a1 - a0 + 31) / 32);
    const int nslots = 2 * npairs + W - 1;

    for (int i = tid; i < TAPS; i += 256) {
        const int t = i - 16;
        tap_lds[i] = (t >= 0 && t <= 2 * RT) ? taps[t] : 0.0f;
    }

    // ---- data loads (as in the one-tile form)
    const int dr = 4 * wv + (lane >> 4);
    const int dl = (4 * (lane & 15) - 16 * (dr & 3)) & 63;
    const bool dl_ok = line0 + dl < g.n_lane;
    const char *dsrc = reinterpret_cast<const char *>(in + outer * g.in_outer + line0 + dl);
    const int64_t in_ax_b = g.in_ax * 4;
    const int64_t pos_first = a0 - RT;
    const uint32_t ring_lds = lds_addr(ring) + 1024u * (uint32_t)wv;
    auto issue_data = [&](int slot) {
        const int64_t p = pos_first + 16 * (int64_t)slot + dr;
        const bool ok = dl_ok && slot < nslots && p >= 0 && p < n_ax;
        const void *src = ok ? static_cast<const void *>(dsrc + p * in_ax_b)
                             : static_cast<const void *>(g_dma_zero);
        dma16<ZSTAGE>(ring_lds + (uint32_t)((slot % NSL) * kF32Slot * 4), src);
    };
    // ---- mask bits, four pairs per load: lane -> pair pj + lane / 16, line
    // 16 wv + lane % 16 (the 32-bit word of its 32 positions; segments start
    // at multiples of 32), so a slot holds [wave][4 pairs][16 lines] and each
    // load reads 64 contiguous bytes per pair
    const int ml = 16 * wv + (lane & 15), mk = lane >> 4;
    const bool ml_ok = MASK && line0 + ml < g.n_lane;
    const uint32_t *msrc = MASK ? e.bits.w + outer * e.bits.outer + (line0 + ml) * e.bits.lane : nullptr;
    const uint32_t mring_lds = lds_addr(mring) + 256u * (uint32_t)wv;
    auto issue_mask = [&](int pj) {
        const int64_t ppos = a0 + 32 * (int64_t)(pj + mk);
        const bool ok = ml_ok && ppos + 32 <= a1;
        const void *src = ok ? static_cast<const void *>(msrc + (ppos / 32) * e.bits.word)
                             : static_cast<const void *>(g_dma_zero);
        dma4(mring_lds + (uint32_t)(((pj >> 2) % NM4) * 1024), src);
    };
    auto issue_group = [&](int i) {
        issue_data(2 * i + W - 1);
        issue_data(2 * i + W);
        if (MASK && (i & 3) == 0) issue_mask(i);
    };

    for (int s = 0; s < W - 1; ++s) issue_data(s);
    for (int i = 0; i < D; ++i) issue_group(i);
    __syncthreads();  // tap table

    float tv[C];
#pragma unroll
    for (int c = 0; c < C; ++c) tv[c] = tap_lds[4 * c + kg - col + 16];
    const int rd_lane = kg * 64 + ((16 * wv + col + 16 * kg) & 63);

    // store-side constants
    const int zr = tid >> 4, zq = tid & 15;  // line-contiguous: positions zr, zr + 16; 4 lines
    const bool zq_ok = line0 + 4 * zq < g.n_lane;
    const int sl = tid >> 2, sq = tid & 3;   // position-contiguous: line, 8 positions
    const bool sl_ok = line0 + sl < g.n_lane;
    float *const sinkf = reinterpret_cast<float *>(g_dma_sink + lane);
    uint16_t *const sinku = reinterpret_cast<uint16_t *>(g_dma_sink + lane);

    for (int j = 0; j < npairs; ++j) {
        if (j >= D)
            vm_wait<GW + D * ST>();
        else
            vm_wait<GW>();
        __syncthreads();
        issue_group(j + D);
        float xs[C + 4];
#pragma unroll
        for (int i = 0; i < C + 4; ++i) {
            int si = 2 * j + i / 4;
            si %= NSL;
            xs[i] = ring[si * kF32Slot + (i % 4) * 256 + rd_lane];
        }
        v4f acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c = 0; c < C; ++c) {
            acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(tv[c], xs[c], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(tv[c], xs[c + 4], acc1, 0, 0, 0);
        }
        const int64_t ppos = a0 + 32 * (int64_t)j;
        const bool full = ppos + 32 <= a1;  // uniform; only a segment's last pair is partial
        if (ZSTAGE) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                stage[(4 * kg + r) * kF32StagePitchZ + 16 * wv + col] = acc0[r];
                stage[(16 + 4 * kg + r) * kF32StagePitchZ + 16 * wv + col] = acc1[r];
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int row = zr + 16 * h;
                const v4f v = *reinterpret_cast<const v4f *>(&stage[row * kF32StagePitchZ + 4 * zq]);
                const int64_t p = ppos + row;
                float *op = e.out_f32 + outer * g.out_outer + p * g.out_ax + line0 + 4 * zq;
                if (full) {
                    if (nt_rows_ok(e.out_f32, 4 * g.out_outer))
                        store16<true>(zq_ok ? op : sinkf, v);
                    else
                        store16<false>(zq_ok ? op : sinkf, v);
                } else if (zq_ok && p < a1) {
                    store16<false>(op, v);
                }
            }
        } else {
            *reinterpret_cast<v4f *>(&stage[(16 * wv + col) * kF32StagePitchV + 4 * kg]) = acc0;
            *reinterpret_cast<v4f *>(&stage[(16 * wv + col) * kF32StagePitchV + 16 + 4 * kg]) = acc1;
            __syncthreads();
            v4f v[2];
#pragma unroll
            for (int h = 0; h < 2; ++h)
                v[h] = *reinterpret_cast<const v4f *>(&stage[sl * kF32StagePitchV + 8 * sq + 4 * h]);
            const int64_t p0 = ppos + 8 * sq;
            const int64_t off = outer * g.out_outer + (line0 + sl) * g.out_lane + p0;
            if (full) {
                if (MASK) {
                    const uint32_t m = *reinterpret_cast<const uint32_t *>(
                        mring + ((j >> 2) % NM4) * 1024 + (sl >> 4) * 256 + (j & 3) * 64 + (sl & 15) * 4) >>
                        (8 * sq);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        if ((m >> r) & 1u) v[0][r] = 1.0f;
                        if ((m >> (4 + r)) & 1u) v[1][r] = 1.0f;
                    }
                }
                if (OUTF) {
                    float *dst = sl_ok ? e.out_f32 + off : sinkf;
                    store16<false>(dst, v[0]);
                    store16<false>(dst + (sl_ok ? 4 : 0), v[1]);
                }
                if (OUTU) {
                    uint32_t w[4];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        w[2 * h] = (uint32_t)quantise_u16_l(v[h][0]) | ((uint32_t)quantise_u16_l(v[h][1]) << 16);
                        w[2 * h + 1] = (uint32_t)quantise_u16_l(v[h][2]) | ((uint32_t)quantise_u16_l(v[h][3]) << 16);
                    }
                    store16<false>(sl_ok ? e.out_u16 + off : sinku, make_uint4(w[0], w[1], w[2], w[3]));
                }
            } else if (sl_ok) {
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    if (p0 + r >= a1) break;
                    float w = v[r >> 2][r & 3];
                    if (MASK && gmem(e.mask)[off + r] != 0) w = 1.0f;
                    if (OUTF) gmem(e.out_f32)[off + r] = w;
                    if (OUTU) gmem(e.out_u16)[off + r] = quantise_u16_l(w);
                }
            }
        }
    }
    vm_wait<0>();
    __syncthreads();
}

// Which body serves a form (measured on 1024^3 and 2048x2048x1536): the
// line-contiguous z-pass is fastest with one tile per iteration and two slots
// in flight; the position-contiguous passes with two tiles per iteration (one
// group in flight), which amortises their staged epilogue over twice the work.
template <int RT, int D, bool MASK, bool ZSTAGE>
struct F32DmaForm {
    static constexpr int bytes() {
        if constexpr (ZSTAGE)
            return F32DmaSmem<RT, D, false, true>::BYTES;  // the line-contiguous pass never illuminates
        else
            return F32Dma2Smem<RT, MASK ? MXB_F32_DEPTH_X : MXB_F32_DEPTH_Y, MASK, false>::BYTES;
    }
    static constexpr int BYTES = bytes();
};

template <int RT, int D, bool MASK, bool OUTF, bool OUTU, bool ZSTAGE, typename TapSrc>
__device__ __forceinline__ void f32_dma_pass(const float *__restrict__ in, const LineGeom &g,
                                             const Epilogue &e, int64_t outer, int64_t lb,
                                             int64_t seg_id, const TapSrc &taps, unsigned char *smem) {
    if constexpr (ZSTAGE)
        f32_dma_body<RT, D, MASK, OUTF, OUTU, ZSTAGE>(in, g, e, outer, lb, seg_id, taps, smem);
    else
        f32_dma_body2<RT, MASK ? MXB_F32_DEPTH_X : MXB_F32_DEPTH_Y, MASK, OUTF, OUTU, ZSTAGE>(in, g, e, outer, lb,
                                                                                           seg_id, taps, smem);
}

}  // namespace mxb
