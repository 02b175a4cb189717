// Direct-load body of the reduced-precision (BF16 / FP16 state) line pass
// (gfx950), shared by the single-volume kernel (mxb_kernels_lowp_lds.hip) and
// the batched persistent kernel (mxb_batched.hip).  Arithmetic, chunking and
// tap split as in mxb_lowp.h; row layout and transposed operand reads as
// described in mxb_kernels_lowp_lds.hip.
#ifndef MXB_LOWP_DMA_BODY_H
#define MXB_LOWP_DMA_BODY_H

#include "mxb_dma.h"
#include "mxb_lowp.h"

// Depth of the direct-load prefetch (tile pairs ahead) of the single-volume
// passes: the line-contiguous z-pass and the position-contiguous y / x passes
// (the masked x-pass needs D - 1 to be a multiple of 4).
#ifndef MXB_LOWP_DEPTH_Z
#define MXB_LOWP_DEPTH_Z 5
#endif
#ifndef MXB_LOWP_DEPTH_YX
#define MXB_LOWP_DEPTH_YX 9
#endif

namespace mxb {

constexpr int kSlotBytes = 32 * 128;       // 32 positions x 64 lines x 2 bytes
constexpr int kStagePitchZ = 144;          // bytes per position row (z-staged epilogue)
constexpr int kStagePitchV = 80;           // bytes per line row (position-contiguous epilogue)
constexpr int kStageBytesL = 64 * kStagePitchV;  // >= 32 * kStagePitchZ

__device__ __forceinline__ int row_swizzle(int r) { return ((r >> 1) & 1) | ((r >> 2) & 2); }

// ---------------------------------------------------------------------------
// Direct-load form: the same pass with the slots (and, for the illuminating
// pass, the 64 lines' 32-bit mask words, four pairs at a time) moved global
// -> LDS by global_load_lds_dword{x4,}, D slots ahead of use, with no VGPR
// staging.
//
// Iteration j (tile pair j):
//   A. wait until group j has landed (vm_wait), barrier;
//   B. issue group j + D into the ring slots freed by pair j - 1;
//   C. compute pair j from slots j .. j + W - 1;
//   D. stage the outputs in LDS, barrier, one 16-byte store per thread.
// Group i = data slot i + W - 1 (+, when i % 4 == 0, the mask words of pairs
// i .. i+3, one 4-byte load per lane).  The prologue issues data slots
// 0 .. W - 2 and groups 0 .. D - 1.  After group j the wave has issued groups
// j+1 .. j+D-1 (D - 1 data loads and, with (D - 1) % 4 == 0, exactly
// (D - 1) / 4 mask loads) and the stores of up to D pairs, so the wait is
// vmcnt(GW + min(j, D) S) with GW those loads and S = 1 store per pair; the
// min(j, D) term is dropped (a smaller count only waits longer) for the
// first D pairs.  Loads outside the volume or past the
// segment read g_dma_zero and stores of lines beyond the volume go to
// g_dma_sink, so every wave issues exactly these counts; the last pair of
// a segment, when partial, is written element by element after its wait.
template <int RT, int D>
struct LowpDmaShape {
    static constexpr int C = LowpChunks<RT>::C;
    static constexpr int W = C + 1;    // slots a tile pair reads
    static constexpr int NSL = W + D;  // ring slots
    static constexpr int NM = (D + 7) / 4;  // mask slots of 4 pairs: reused only after pair i+3's epilogue
};

constexpr int kMaskSlotBytes = 64 * 16;  // 64 lines x 4 pairs' 32-bit words of mask bits

constexpr int kStagePitchF = 36;  // floats per line row (position-contiguous FP32 epilogue)

// SPLIT: two rings (high and low planes) and two 16-bit stages; FIN != 0: the
// position-contiguous epilogue stages floats; WIDEN != 0 (split form): one
// ring, the input being one 16-bit plane (1 BF16, 2 FP16) taken as the high
// plane with a zero low plane.
template <int RT, int D, bool MASK, bool ZSTAGE, bool SPLIT = false, int FIN = 0, int WIDEN = 0>
struct LowpDmaSmem {
    using S = LowpDmaShape<RT, D>;
    static constexpr int RING = S::NSL * kSlotBytes * (SPLIT && WIDEN == 0 ? 2 : 1);
    static constexpr int MRING = MASK ? S::NM * kMaskSlotBytes : 0;
    static constexpr int STAGE = FIN != 0 ? 64 * kStagePitchF * 4 : kStageBytesL * (SPLIT ? 2 : 1);
    static constexpr int TAPS = 2 * LowpChunks<RT>::TW * 2;  // hi and lo tables, uint16
    static constexpr int BYTES = RING + MRING + STAGE + TAPS;
};

// One (outer, 64-line group lb, segment) unit of a pass; `smem` holds
// LowpDmaSmem::BYTES of LDS, `tap_hi` / `tap_lo` the 2 RT + 1 centred taps'
// 16-bit high and low parts.  Ends with every load and store drained and a
// barrier, so a persistent caller may reuse `smem` at once.
template <int RT, int D, bool BF, bool MASK, bool ZSTAGE, bool SPLIT = false, int FIN = 0, int WIDEN = 0>
__device__ __forceinline__ void lowp_dma_body(const uint16_t *__restrict__ in, uint16_t *__restrict__ out,
                                              const uint8_t *__restrict__ mask, const MaskBits &bits,
                                              const LineGeom &g, const uint16_t *tap_hi,
                                              const uint16_t *tap_lo, int64_t outer, int64_t lb,
                                              int64_t seg_id, unsigned char *smem,
                                              const SplitIO &io = SplitIO{}) {
    using S = LowpDmaShape<RT, D>;
    using L = Lowp<BF>;
    constexpr int C = S::C, W = S::W, NSL = S::NSL, NM = S::NM;
    constexpr int TW = LowpChunks<RT>::TW;
    constexpr bool LO = SPLIT && WIDEN == 0;  // a low plane is read
    constexpr int P = SPLIT ? 2 : 1;          // state planes written
    constexpr int GW = (LO ? 2 : 1) * (D - 1) + (MASK ? (D - 1) / 4 : 0);  // loads in D - 1 groups
    // stores per pair and thread
    constexpr int ST = FIN != 0 ? ((FIN & 1) ? 2 : 0) + ((FIN & 2) ? 1 : 0) : P;
    static_assert(!MASK || (D - 1) % 4 == 0, "a fixed count of mask loads per D - 1 groups");
    static_assert(!(MASK && ZSTAGE), "the illuminating pass writes positions contiguously");
    static_assert(!SPLIT || !BF, "the split form runs on FP16 MFMAs");
    static_assert(FIN == 0 || (SPLIT && !ZSTAGE), "final outputs: the split form's position-contiguous pass");
    static_assert(WIDEN == 0 || (SPLIT && ZSTAGE && !MASK), "widening: the split form's z-pass");
    using M = LowpDmaSmem<RT, D, MASK, ZSTAGE, SPLIT, FIN, WIDEN>;
    unsigned char *const ring = smem;
    unsigned char *const ring_lo = smem + S::NSL * kSlotBytes;  // SPLIT only
    unsigned char *const mring = smem + M::RING;
    unsigned char *const stage = smem + M::RING + M::MRING;
    uint16_t *const thi = reinterpret_cast<uint16_t *>(smem + M::RING + M::MRING + M::STAGE);
    uint16_t *const tlo = thi + TW;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int col = lane & 15, kg = lane >> 4;
    const int64_t line0 = lb * 64;
    const int64_t a0 = seg_id * g.seg;
    const int64_t a1 = a0 + g.seg < g.n_ax ? a0 + g.seg : g.n_ax;
    const int64_t n_ax = g.n_ax;
    const int npairs = (int)((a1 - a0 + 31) / 32);
    const int nslots = npairs + W - 1;  // slots the segment reads

    for (int i = threadIdx.x; i < TW; i += blockDim.x) {
        const int t = i - 16;
        const bool in_band = t >= 0 && t <= 2 * RT;
        thi[i] = in_band ? tap_hi[in_band ? t : 0] : 0;
        tlo[i] = in_band ? tap_lo[in_band ? t : 0] : 0;
    }

    // ---- data loads: lane -> row dr of the slot, 8 lines from dl (swizzled quads)
    const int dr = 8 * wv + (lane >> 3);
    const int pp = lane & 7;
    const int dl = 16 * ((pp >> 1) ^ row_swizzle(dr)) + 8 * (pp & 1);
    const bool dl_ok = line0 + dl < g.n_lane;  // n_lane is a multiple of 8
    const char *dsrc = reinterpret_cast<const char *>(in + outer * g.in_outer + line0 + dl);
    const int64_t in_ax_b = g.in_ax * 2;
    const int64_t pos_first = a0 - RT;  // position of slot 0, row 0
    const uint32_t ring_lds = lds_addr(ring) + 1024u * (uint32_t)wv;
    const int64_t in_lo_b = io.in_lo * 2;
    auto issue_data = [&](int slot) {
        const int64_t p = pos_first + 32 * (int64_t)slot + dr;
        const bool ok = dl_ok && slot < nslots && p >= 0 && p < n_ax;
        const char *src = ok ? dsrc + p * in_ax_b : reinterpret_cast<const char *>(g_dma_zero);
        dma16<ZSTAGE>(ring_lds + (uint32_t)((slot % NSL) * kSlotBytes), src);
        if constexpr (LO)
            dma16<ZSTAGE>(ring_lds + (uint32_t)((NSL + slot % NSL) * kSlotBytes), ok ? src + in_lo_b : src);
    };
    // ---- mask bits: lane -> pair pj + lane / 16, line 16 wv + lane % 16 (the
    // 32-bit word of its 32 positions; segments start at multiples of 32), so
    // a slot holds [wave][4 pairs][16 lines] and each load reads 64
    // contiguous bytes per pair
    const int ml = 16 * wv + (lane & 15), mk = lane >> 4;
    const bool ml_ok = MASK && line0 + ml < g.n_lane;
    const uint32_t *msrc = MASK ? bits.w + outer * bits.outer + (line0 + ml) * bits.lane : nullptr;
    const uint32_t mring_lds = lds_addr(mring) + 256u * (uint32_t)wv;
    auto issue_mask = [&](int pj) {
        const int64_t ppos = a0 + 32 * (int64_t)(pj + mk);
        const bool ok = ml_ok && ppos + 32 <= a1;
        const void *src = ok ? static_cast<const void *>(msrc + (ppos / 32) * bits.word)
                             : static_cast<const void *>(g_dma_zero);
        dma4(mring_lds + (uint32_t)(((pj >> 2) % NM) * kMaskSlotBytes), src);
    };
    auto issue_group = [&](int i) {
        issue_data(i + W - 1);
        if (MASK && (i & 3) == 0) issue_mask(i);
    };

    for (int s = 0; s < W - 1; ++s) issue_data(s);
    for (int i = 0; i < D; ++i) issue_group(i);
    __syncthreads();  // tap tables

    v8s ahi[C], alo[C];
    lowp_read_taps<RT>(thi, tlo, col, kg, ahi, alo);

    int roff[2][2], radj[2][2];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int rl = 16 * u + 8 * kg + (col >> 2) + 4 * h;
            const int rr = rl & 31;
            radj[u][h] = rl >> 5;
            roff[u][h] = rr * 128 + ((wv ^ row_swizzle(rr)) << 5) + 8 * (col & 3);
        }

    // store-side constants
    const int sl = tid >> 2, sq = tid & 3;         // position-contiguous: line, 8 positions
    const int zr = tid >> 3, zq = tid & 7;         // line-contiguous: position, 8 lines
    const bool sl_ok = line0 + sl < g.n_lane;
    const bool zq_ok = line0 + 8 * zq < g.n_lane;
    uint16_t *const sink = reinterpret_cast<uint16_t *>(g_dma_sink + lane);

    for (int j = 0; j < npairs; ++j) {
        // A. group j has landed in every wave
        if (j >= D)
            vm_wait<GW + D * ST>();
        else
            vm_wait<GW>();
        __syncthreads();
        // B. prefetch group j + D
        issue_group(j + D);
        // C. tiles 2j, 2j+1 from slots j .. j + W - 1
        const int sj = j % NSL;
        v4f acc[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[u] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                typedef __attribute__((address_space(3))) v4s lds_v4s;
                v4s part[2], plo[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    int si = sj + c + radj[u][h];
                    si = si >= NSL ? si - NSL : si;
                    part[h] = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
                        (lds_v4s *)(ring + si * kSlotBytes + roff[u][h]));
                    if constexpr (LO)
                        plo[h] = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
                            (lds_v4s *)(ring_lo + si * kSlotBytes + roff[u][h]));
                }
                v8s x = {part[0][0], part[0][1], part[0][2], part[0][3],
                         part[1][0], part[1][1], part[1][2], part[1][3]};
                if constexpr (WIDEN == 1) x = bf16x8_to_f16(x);
                acc[u] = L::mfma(ahi[c], x, acc[u]);
                acc[u] = L::mfma(alo[c], x, acc[u]);
                if constexpr (LO) {
                    const v8s xl = {plo[0][0], plo[0][1], plo[0][2], plo[0][3],
                                    plo[1][0], plo[1][1], plo[1][2], plo[1][3]};
                    acc[u] = L::mfma(ahi[c], xl, acc[u]);
                }
            }
        // D. epilogue
        const int64_t ppos = a0 + 32 * (int64_t)j;
        const bool full = ppos + 32 <= a1;  // uniform; only a segment's last pair is partial
        if constexpr (ZSTAGE && !SPLIT) {
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    *reinterpret_cast<uint16_t *>(stage + (16 * u + 4 * kg + r) * kStagePitchZ +
                                                  (16 * wv + col) * 2) = L::from_f32(acc[u][r]);
            __syncthreads();
            const uint4 v = *reinterpret_cast<const uint4 *>(stage + zr * kStagePitchZ + 16 * zq);
            const int64_t p = ppos + zr;
            uint16_t *op = out + outer * g.out_outer + p * g.out_ax + line0 + 8 * zq;
            if (full) {
                if (nt_rows_ok(out, 2 * g.out_outer))
                    store16<true>(zq_ok ? op : sink, v);
                else
                    store16<false>(zq_ok ? op : sink, v);
            } else if (zq_ok && p < a1) {
                *reinterpret_cast<uint4 *>(op) = v;
            }
        } else if constexpr (ZSTAGE) {
            // split planes, line-contiguous
            unsigned char *const stage_lo = stage + kStageBytesL;
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    uint16_t hi, lo;
                    split_f16_scaled(acc[u][r], io.scale, hi, lo);
                    const int o = (16 * u + 4 * kg + r) * kStagePitchZ + (16 * wv + col) * 2;
                    *reinterpret_cast<uint16_t *>(stage + o) = hi;
                    *reinterpret_cast<uint16_t *>(stage_lo + o) = lo;
                }
            __syncthreads();
            const uint4 vh = *reinterpret_cast<const uint4 *>(stage + zr * kStagePitchZ + 16 * zq);
            const uint4 vl = *reinterpret_cast<const uint4 *>(stage_lo + zr * kStagePitchZ + 16 * zq);
            const int64_t p = ppos + zr;
            uint16_t *op = out + outer * g.out_outer + p * g.out_ax + line0 + 8 * zq;
            if (full) {
                if (nt_rows_ok(out, 2 * g.out_outer)) {
                    store16<true>(zq_ok ? op : sink, vh);
                    store16<true>(zq_ok ? op + io.out_lo : sink, vl);
                } else {
                    store16<false>(zq_ok ? op : sink, vh);
                    store16<false>(zq_ok ? op + io.out_lo : sink, vl);
                }
            } else if (zq_ok && p < a1) {
                *reinterpret_cast<uint4 *>(op) = vh;
                *reinterpret_cast<uint4 *>(op + io.out_lo) = vl;
            }
        } else if constexpr (FIN != 0) {
            // split form, final pass: FP32 and / or uint16 outputs
            float *const stagef = reinterpret_cast<float *>(stage);
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const v4f a = acc[u] * io.scale;
                *reinterpret_cast<v4f *>(stagef + (16 * wv + col) * kStagePitchF + 16 * u + 4 * kg) = a;
            }
            __syncthreads();
            v4f v[2];
#pragma unroll
            for (int h = 0; h < 2; ++h)
                v[h] = *reinterpret_cast<const v4f *>(stagef + sl * kStagePitchF + 8 * sq + 4 * h);
            const int64_t p0 = ppos + 8 * sq;
            const int64_t off = outer * g.out_outer + (line0 + sl) * g.out_lane + p0;
            float *const sinkf = reinterpret_cast<float *>(g_dma_sink + lane);
            if (full) {
                if (MASK) {
                    const uint32_t m = *reinterpret_cast<const uint32_t *>(
                        mring + ((j >> 2) % NM) * kMaskSlotBytes + (sl >> 4) * 256 + (j & 3) * 64 + (sl & 15) * 4) >>
                        (8 * sq);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        if ((m >> r) & 1u) v[0][r] = 1.0f;
                        if ((m >> (4 + r)) & 1u) v[1][r] = 1.0f;
                    }
                }
                if constexpr ((FIN & 1) != 0) {
                    float *dst = sl_ok ? io.fin_f32 + off : sinkf;
                    store16<false>(dst, v[0]);
                    store16<false>(dst + (sl_ok ? 4 : 0), v[1]);
                }
                if constexpr ((FIN & 2) != 0) {
                    uint32_t w[4];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        w[2 * h] = (uint32_t)quantise_u16_s(v[h][0]) | ((uint32_t)quantise_u16_s(v[h][1]) << 16);
                        w[2 * h + 1] = (uint32_t)quantise_u16_s(v[h][2]) | ((uint32_t)quantise_u16_s(v[h][3]) << 16);
                    }
                    store16<false>(sl_ok ? io.fin_u16 + off : sink, make_uint4(w[0], w[1], w[2], w[3]));
                }
            } else if (sl_ok) {
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    if (p0 + r >= a1) break;
                    float w = v[r >> 2][r & 3];
                    if (MASK && gmem(mask)[off + r] != 0) w = 1.0f;
                    if constexpr ((FIN & 1) != 0) gmem(io.fin_f32)[off + r] = w;
                    if constexpr ((FIN & 2) != 0) gmem(io.fin_u16)[off + r] = quantise_u16_s(w);
                }
            }
        } else {
            // position-contiguous 16-bit planes (one, or the split form's two)
            unsigned char *const stage_lo = stage + kStageBytesL;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if constexpr (SPLIT) {
                    uint16_t h[4], l[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) split_f16_scaled(acc[u][r], io.scale, h[r], l[r]);
                    const int o = (16 * wv + col) * kStagePitchV + (16 * u + 4 * kg) * 2;
                    *reinterpret_cast<uint2 *>(stage + o) = make_uint2(pack2(h[0], h[1]), pack2(h[2], h[3]));
                    *reinterpret_cast<uint2 *>(stage_lo + o) = make_uint2(pack2(l[0], l[1]), pack2(l[2], l[3]));
                } else {
                    const uint2 v = make_uint2(pack2(L::from_f32(acc[u][0]), L::from_f32(acc[u][1])),
                                               pack2(L::from_f32(acc[u][2]), L::from_f32(acc[u][3])));
                    *reinterpret_cast<uint2 *>(stage + (16 * wv + col) * kStagePitchV + (16 * u + 4 * kg) * 2) = v;
                }
            }
            __syncthreads();
            uint4 v = *reinterpret_cast<const uint4 *>(stage + sl * kStagePitchV + 16 * sq);
            uint4 vl = make_uint4(0, 0, 0, 0);
            if constexpr (SPLIT) vl = *reinterpret_cast<const uint4 *>(stage_lo + sl * kStagePitchV + 16 * sq);
            const int64_t p0 = ppos + 8 * sq;
            const int64_t off = outer * g.out_outer + (line0 + sl) * g.out_lane + p0;
            if (full) {
                if (MASK) {
                    const uint32_t m = *reinterpret_cast<const uint32_t *>(
                        mring + ((j >> 2) % NM) * kMaskSlotBytes + (sl >> 4) * 256 + (j & 3) * 64 + (sl & 15) * 4) >>
                        (8 * sq);
                    const uint32_t one = L::from_f32(1.0f);
                    uint32_t w[4] = {v.x, v.y, v.z, v.w};
                    uint32_t wl[4] = {vl.x, vl.y, vl.z, vl.w};
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        if ((m >> e) & 1u) {
                            const int sh = 16 * (e & 1);
                            w[e >> 1] = (w[e >> 1] & ~(0xffffu << sh)) | (one << sh);
                            wl[e >> 1] &= ~(0xffffu << sh);
                        }
                    }
                    v = make_uint4(w[0], w[1], w[2], w[3]);
                    vl = make_uint4(wl[0], wl[1], wl[2], wl[3]);
                }
                store16<false>(sl_ok ? out + off : sink, v);
                if constexpr (SPLIT) store16<false>(sl_ok ? out + io.out_lo + off : sink, vl);
            } else if (sl_ok && p0 < a1) {
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
                const uint32_t wl[4] = {vl.x, vl.y, vl.z, vl.w};
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    if (p0 + e >= a1) break;
                    uint16_t b = (uint16_t)(w[e >> 1] >> (16 * (e & 1)));
                    uint16_t bl = (uint16_t)(wl[e >> 1] >> (16 * (e & 1)));
                    if (MASK && mask[off + e] != 0) {
                        b = L::from_f32(1.0f);
                        bl = 0;
                    }
                    out[off + e] = b;
                    if constexpr (SPLIT) out[io.out_lo + off + e] = bl;
                }
            }
        }
    }
    vm_wait<0>();
    __syncthreads();
}

}  // namespace mxb

#endif  // MXB_LOWP_DMA_BODY_H
