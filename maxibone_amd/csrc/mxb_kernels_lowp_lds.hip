// Reduced-precision (BF16 / FP16 state) line pass, LDS-staged form (gfx950).
//
// Arithmetic, chunking and tap split as in mxb_lowp.h (the register-ring
// kernel computes the same values bit for bit).  The difference is how the
// MFMA B operands (8 consecutive positions of one line per lane) are formed:
//
// * global -> VGPR: one 16-byte load per thread moves a slot of 32 positions
//   x 64 lines (4 KB, 32 full 128-byte rows), one row piece of 8 lines each;
// * VGPR -> LDS: the slot is stored position-major (a row of 64 lines per
//   position) into a ring of NSL slots;
// * LDS -> VGPR: ds_read_b64_tr_b16 transposes on the way out.  Within each
//   16-lane group, lane i receives element i % 4 of the 8-byte pieces that
//   lanes i / 4 + 4k (k = 0..3) address.  Lane q of group g addresses row
//   p0 + 8g + q / 4, lines 4 (q % 4) .. +3 of the wave's 16 lines, so lane
//   i gets positions p0 + 8g .. +3 of line i; a second read at p0 + 4 gives
//   the other half of the 8-position operand.
//
// Block = 256 threads = 4 waves over 64 consecutive lines; wave w computes
// lines 16w .. 16w+15.  One iteration computes a pair of 16-output tiles
// (32 positions) per wave from W = C + 1 slots, while the next slot is being
// written to the ring from the VGPRs loaded one iteration earlier and the
// global loads for the slot after it are in flight.
//
// Row layout: the 128-byte row of position r holds its four 16-line quads
// in quad order w ^ s(r), s(r) = bit1(r) | bit3(r) << 1.  The rows one
// transposed read touches are p0 + {0..3, 8..11, 16..19, 24..27} with
// p0 = 0 mod 4; per half-wave the eight 32-byte pieces then fall in eight
// distinct 32-byte bank groups.
#include "maxibone_diffusion.h"
#include "mxb_dma.h"
#include "mxb_lowp.h"
#include "mxb_lowp_dma_body.h"

namespace mxb {

template <int RT>
struct LowpLdsShape {
    static constexpr int C = LowpChunks<RT>::C;
    static constexpr int W = C + 1;      // slots a tile pair reads (16 + 32C rows)
    static constexpr int NSL = W + 1;    // ring slots
};


template <int RT, bool BF>
__global__ __launch_bounds__(256) void k_line_lowp_lds(const uint16_t *__restrict__ in,
                                                       uint16_t *__restrict__ out,
                                                       const uint8_t *__restrict__ mask,
                                                       LineGeom g, TapsL<2 * RT + 1> taps) {
    using S = LowpLdsShape<RT>;
    using L = Lowp<BF>;
    constexpr int C = S::C, W = S::W, NSL = S::NSL;
    constexpr int TW = LowpChunks<RT>::TW;
    __shared__ __attribute__((aligned(16))) unsigned char ring[NSL * kSlotBytes];
    __shared__ __attribute__((aligned(16))) unsigned char stage[kStageBytesL];
    __shared__ uint16_t thi[TW], tlo[TW];

    int64_t bid = blockIdx.x;
    const int64_t seg_id = bid % g.nseg;
    bid /= g.nseg;
    const int64_t nlb = (g.n_lane + 63) >> 6;
    const int64_t lb = bid % nlb;
    const int64_t outer = bid / nlb;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int col = lane & 15, kg = lane >> 4;
    const int64_t line0 = lb * 64;
    const int64_t a0 = seg_id * g.seg;
    const int64_t a1 = a0 + g.seg < g.n_ax ? a0 + g.seg : g.n_ax;
    const int64_t n_ax = g.n_ax;

    lowp_fill_taps<RT>(thi, tlo, taps);

    // ---- loader: thread -> (row lr = tid / 8, lines 8 (tid % 8) .. +7)
    const int lr = tid >> 3;
    const int lq = tid & 7;
    const int64_t lline = line0 + 8 * lq;
    // lines beyond n_lane read line n_lane-1 (their outputs are never stored)
    const bool oct_full = lline + 7 < g.n_lane;
    const char *lbase = reinterpret_cast<const char *>(in + outer * g.in_outer);
    const int64_t in_ax_b = g.in_ax * 2;
    const int64_t pos_first = a0 - RT;  // position of slot 0, row 0
    auto gload = [&](int64_t slot) -> uint4 {
        const int64_t p = pos_first + 32 * slot + lr;
        if (p < 0 || p >= n_ax) return make_uint4(0u, 0u, 0u, 0u);
        const char *rowp = lbase + p * in_ax_b;
        if (oct_full) return *reinterpret_cast<const uint4 *>(rowp + lline * 2);
        uint16_t v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int64_t l = lline + q < g.n_lane ? lline + q : g.n_lane - 1;
            v[q] = *reinterpret_cast<const uint16_t *>(rowp + l * 2);
        }
        return make_uint4(pack2(v[0], v[1]), pack2(v[2], v[3]), pack2(v[4], v[5]), pack2(v[6], v[7]));
    };
    const int woff = lr * 128 + (((lq >> 1) ^ row_swizzle(lr)) << 5) + 16 * (lq & 1);
    auto lwrite = [&](int64_t slot, uint4 v) {
        *reinterpret_cast<uint4 *>(ring + (int)(slot % NSL) * kSlotBytes + woff) = v;
    };

    for (int s = 0; s < W; ++s) lwrite(s, gload(s));
    uint4 staged = gload(W);
    __syncthreads();

    v8s ahi[C], alo[C];
    lowp_read_taps<RT>(thi, tlo, col, kg, ahi, alo);

    // ---- transposed-read offsets: tile u of the pair, half h of the operand.
    // Row (relative to the pair's first slot) rl = 16u + 8kg + col/4 + 4h
    // (+ 32c for chunk c): slot adj = rl / 32, row rl % 32.
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

    const int64_t li = line0 + 16 * wv + col;
    const bool act = li < g.n_lane;
    const int64_t npairs = (a1 - a0 + 31) / 32;

    for (int64_t j = 0; j < npairs; ++j) {
        // 1. slot j + W (loaded last iteration) into the ring slot freed by pair j-1
        lwrite(j + W, staged);
        // 2. global loads for slot j + W + 1 (written next iteration)
        if (j + 1 < npairs) staged = gload(j + W + 1);
        // 3. tiles 2j, 2j+1 from slots j .. j + W - 1
        const int sj = (int)(j % NSL);
        v4f acc[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[u] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                v4s part[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    int si = sj + c + radj[u][h];
                    si = si >= NSL ? si - NSL : si;
                    typedef __attribute__((address_space(3))) v4s lds_v4s;
                    part[h] = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
                        (lds_v4s *)(ring + si * kSlotBytes + roff[u][h]));
                }
                const v8s x = {part[0][0], part[0][1], part[0][2], part[0][3],
                               part[1][0], part[1][1], part[1][2], part[1][3]};
                acc[u] = L::mfma(ahi[c], x, acc[u]);
                acc[u] = L::mfma(alo[c], x, acc[u]);
            }
        // 4. epilogue: tile u holds outputs pair_pos + 16u + 4kg + r of line li
        const int64_t ppos = a0 + 32 * j;
        if (g.stage_out) {
            // lines contiguous in the output: stage 32 positions x 64 lines,
            // then every thread stores 8 lines of one position
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    *reinterpret_cast<uint16_t *>(stage + (16 * u + 4 * kg + r) * kStagePitchZ +
                                                  (16 * wv + col) * 2) = L::from_f32(acc[u][r]);
            __syncthreads();
            const int64_t p = ppos + lr;
            if (p < a1) {
                const uint4 v = *reinterpret_cast<const uint4 *>(stage + lr * kStagePitchZ + 16 * lq);
                uint16_t *op = out + outer * g.out_outer + p * g.out_ax;
                if (oct_full) {
                    *reinterpret_cast<uint4 *>(op + lline) = v;
                } else {
                    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (lline + q < g.n_lane) op[lline + q] = (uint16_t)(w[q >> 1] >> (16 * (q & 1)));
                }
            }
        } else if (g.vec_out) {
            // positions contiguous in the output: stage 64 lines x 32
            // positions, then every thread stores 8 positions of one line
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const uint2 v = make_uint2(pack2(L::from_f32(acc[u][0]), L::from_f32(acc[u][1])),
                                           pack2(L::from_f32(acc[u][2]), L::from_f32(acc[u][3])));
                *reinterpret_cast<uint2 *>(stage + (16 * wv + col) * kStagePitchV + (16 * u + 4 * kg) * 2) = v;
            }
            __syncthreads();
            const int line = tid >> 2;
            const int q = tid & 3;
            const int64_t gl = line0 + line;
            const int64_t p0 = ppos + 8 * q;
            if (gl < g.n_lane && p0 < a1) {
                uint4 v = *reinterpret_cast<const uint4 *>(stage + line * kStagePitchV + 16 * q);
                const int64_t off = outer * g.out_outer + gl * g.out_lane + p0;
                if (p0 + 7 < a1) {
                    if (mask != nullptr) {
                        const uint2 m = *reinterpret_cast<const uint2 *>(mask + off);
                        const uint32_t one = L::from_f32(1.0f);
                        uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const uint32_t mb = ((e < 4 ? m.x : m.y) >> (8 * (e & 3))) & 0xffu;
                            if (mb != 0) {
                                const int sh = 16 * (e & 1);
                                w[e >> 1] = (w[e >> 1] & ~(0xffffu << sh)) | (one << sh);
                            }
                        }
                        v = make_uint4(w[0], w[1], w[2], w[3]);
                    }
                    *reinterpret_cast<uint4 *>(out + off) = v;
                } else {
                    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        if (p0 + e >= a1) break;
                        uint16_t b = (uint16_t)(w[e >> 1] >> (16 * (e & 1)));
                        if (mask != nullptr && mask[off + e] != 0) b = L::from_f32(1.0f);
                        out[off + e] = b;
                    }
                }
            }
        } else if (act) {
            const int64_t obase = outer * g.out_outer + li * g.out_lane;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int64_t p0 = ppos + 16 * u + 4 * kg;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (p0 + r >= a1) break;
                    const int64_t off = obase + (p0 + r) * g.out_ax;
                    float v = acc[u][r];
                    if (mask != nullptr && mask[off] != 0) v = 1.0f;
                    out[off] = L::from_f32(v);
                }
            }
        }
        // all waves are done with slot j (and with `stage`) before reuse
        __syncthreads();
    }
}


// ---------------------------------------------------------------------------
// Direct-load form: the body is lowp_dma_body (mxb_lowp_dma_body.h), shared
// with the batched persistent kernel; one block per (segment, 64 lines, outer).
template <int RT, int D, bool BF, bool MASK, bool ZSTAGE>
__global__ __launch_bounds__(256) void k_line_lowp_dma(const uint16_t *__restrict__ in,
                                                       uint16_t *__restrict__ out,
                                                       const uint8_t *__restrict__ mask,
                                                       MaskBits bits, LineGeom g,
                                                       TapsL<2 * RT + 1> taps) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[LowpDmaSmem<RT, D, MASK, ZSTAGE>::BYTES];
    int64_t bid = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t seg_id = bid % g.nseg;
    bid /= g.nseg;
    const int64_t nlb = (g.n_lane + 63) >> 6;
    lowp_dma_body<RT, D, BF, MASK, ZSTAGE>(in, out, mask, bits, g, taps.hi, taps.lo, bid / nlb, bid % nlb,
                                           seg_id, smem);
}

// ---------------------------------------------------------------- host side
template <int RT>
static hipError_t launch_lowp_lds_rt(const uint16_t *in, uint16_t *out, const uint8_t *mask,
                                     LineGeom g, const float *k, int R, bool bf, hipStream_t s) {
    // one block per (segment, 64 lines, outer); segments keep >= ~24 waves/CU
    const int64_t groups = ((g.n_lane + 63) / 64) * g.n_outer;
    const int64_t target = 256LL * 6;
    int64_t nseg = groups > 0 ? (target + groups - 1) / groups : 1;
    const int64_t max_seg = (g.n_ax + 63) / 64;
    if (nseg > max_seg) nseg = max_seg;
    if (nseg < 1) nseg = 1;
    int64_t seg = (g.n_ax + nseg - 1) / nseg;
    seg = ((seg + 31) / 32) * 32;
    g.seg = seg;
    g.nseg = (g.n_ax + seg - 1) / seg;
    g.vec_out = (g.out_ax == 1 && g.out_lane % 8 == 0 && g.out_outer % 8 == 0 &&
                 reinterpret_cast<uintptr_t>(out) % 16 == 0 &&
                 (mask == nullptr || reinterpret_cast<uintptr_t>(mask) % 8 == 0)) ? 1 : 0;
    g.stage_out = (g.out_lane == 1 && g.out_outer % 8 == 0 && g.out_ax % 8 == 0 &&
                   mask == nullptr && reinterpret_cast<uintptr_t>(out) % 16 == 0) ? 1 : 0;
    const int64_t blocks = g.nseg * groups;
    if (blocks == 0) return hipSuccess;
    const TapsL<2 * RT + 1> t = split_taps<RT>(k, R, bf);
    if (bf)
        hipLaunchKernelGGL((k_line_lowp_lds<RT, true>), dim3((unsigned)blocks), dim3(256), 0, s, in, out, mask, g, t);
    else
        hipLaunchKernelGGL((k_line_lowp_lds<RT, false>), dim3((unsigned)blocks), dim3(256), 0, s, in, out, mask, g, t);
    return hipGetLastError();
}



// Direct-load form: needs whole 8-line row pieces (n_lane a multiple of 8) and
// 16-byte-aligned output rows written either line-contiguous or
// position-contiguous (the illuminating pass reads the mask as bits).
static int lowp_dma_form(const uint16_t *out, const uint8_t *mask, const MaskBits &bits, const LineGeom &g) {
    if (g.n_lane % 8 != 0 || reinterpret_cast<uintptr_t>(out) % 16 != 0) return 0;
    if (mask == nullptr && g.out_lane == 1 && g.out_outer % 8 == 0 && g.out_ax % 8 == 0) return 1;
    if (g.out_ax == 1 && g.out_lane % 8 == 0 && g.out_outer % 8 == 0) {
        if (mask == nullptr) return 2;
        if (bits.w != nullptr) return 3;
    }
    return 0;
}

template <int RT, bool BF>
static hipError_t launch_lowp_dma_rt(int form, const uint16_t *in, uint16_t *out, const uint8_t *mask,
                                     const MaskBits &bits, LineGeom g, const float *k, int R,
                                     hipStream_t s) {
    const int64_t groups = ((g.n_lane + 63) / 64) * g.n_outer;
    const int64_t target = 256LL * 6;
    int64_t nseg = groups > 0 ? (target + groups - 1) / groups : 1;
    const int64_t max_seg = (g.n_ax + 63) / 64;
    if (nseg > max_seg) nseg = max_seg;
    if (nseg < 1) nseg = 1;
    int64_t seg = (g.n_ax + nseg - 1) / nseg;
    seg = ((seg + 31) / 32) * 32;
    g.seg = seg;
    g.nseg = (g.n_ax + seg - 1) / seg;
    const int64_t blocks = g.nseg * groups;
    if (blocks == 0) return hipSuccess;
    const TapsL<2 * RT + 1> t = split_taps<RT>(k, R, BF);
    // prefetch depth per form (measured on 1024^3, R=40): the line-contiguous
    // z-pass 5 tile pairs ahead, the position-contiguous passes 9
    if (form == 1)
        hipLaunchKernelGGL((k_line_lowp_dma<RT, MXB_LOWP_DEPTH_Z, BF, false, true>), dim3((unsigned)blocks), dim3(256), 0, s, in, out, mask, bits, g, t);
    else if (form == 2)
        hipLaunchKernelGGL((k_line_lowp_dma<RT, MXB_LOWP_DEPTH_YX, BF, false, false>), dim3((unsigned)blocks), dim3(256), 0, s, in, out, mask, bits, g, t);
    else
        hipLaunchKernelGGL((k_line_lowp_dma<RT, MXB_LOWP_DEPTH_YX, BF, true, false>), dim3((unsigned)blocks), dim3(256), 0, s, in, out, mask, bits, g, t);
    return hipGetLastError();
}

template <int RT>
static hipError_t launch_lowp_staged_rt(const uint16_t *in, uint16_t *out, const uint8_t *mask,
                                        const MaskBits &bits, const LineGeom &g, const float *k, int R,
                                        bool bf, hipStream_t s) {
    const int form = lowp_dma_form(out, mask, bits, g);
    if (form != 0 && !force_vgpr_staging()) {
        if (bf) return launch_lowp_dma_rt<RT, true>(form, in, out, mask, bits, g, k, R, s);
        return launch_lowp_dma_rt<RT, false>(form, in, out, mask, bits, g, k, R, s);
    }
    return launch_lowp_lds_rt<RT>(in, out, mask, g, k, R, bf, s);
}

// ---------------------------------------------------------------- split form
// FP32-class state as FP16 hi + lo planes on the direct-load body (SplitIO in
// mxb_lowp_dma_body.h).  Prefetch depths (tile pairs ahead): the two rings
// double the LDS of a slot, so the depths are lower than the 16-bit forms';
// the illuminating pass needs D - 1 to be a multiple of 4.
#ifndef MXB_SPLIT_DEPTH_Z
#define MXB_SPLIT_DEPTH_Z 3
#endif
#ifndef MXB_SPLIT_DEPTH_Y
#define MXB_SPLIT_DEPTH_Y 3
#endif
#ifndef MXB_SPLIT_DEPTH_X
#define MXB_SPLIT_DEPTH_X 5
#endif

template <int RT, int D, bool MASK, bool ZSTAGE, int FIN, int WIDEN = 0>
__global__ __launch_bounds__(256) void k_line_split_dma(const uint16_t *__restrict__ in,
                                                        uint16_t *__restrict__ out,
                                                        const uint8_t *__restrict__ mask,
                                                        MaskBits bits, LineGeom g,
                                                        TapsL<2 * RT + 1> taps, SplitIO io) {
    __shared__ __attribute__((aligned(16)))
    unsigned char smem[LowpDmaSmem<RT, D, MASK, ZSTAGE, true, FIN, WIDEN>::BYTES];
    int64_t bid = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t seg_id = bid % g.nseg;
    bid /= g.nseg;
    const int64_t nlb = (g.n_lane + 63) >> 6;
    lowp_dma_body<RT, D, false, MASK, ZSTAGE, true, FIN, WIDEN>(in, out, mask, bits, g, taps.hi, taps.lo,
                                                                bid / nlb, bid % nlb, seg_id, smem, io);
}

// Which split-form kernel serves a pass: 0 none; 1 line-contiguous (no
// mask, planes out); 2 position-contiguous planes out; 3 the same,
// illuminating from mask bits; 4 + fin (fin = 1 FP32, 2 uint16, 3 both):
// position-contiguous final outputs; 8 / 9 form 1 from one BF16 / FP16
// plane (io.widen).
static int split_form(const uint16_t *in, const uint16_t *out, const uint8_t *mask, const MaskBits &bits,
                      const LineGeom &g, const SplitIO &io) {
    if (g.n_lane % 8 != 0 || g.in_outer % 8 != 0 || g.in_ax % 8 != 0 || io.in_lo % 8 != 0 ||
        reinterpret_cast<uintptr_t>(in) % 16 != 0)
        return 0;
    if (io.widen != 0) {
        const bool line = mask == nullptr && io.fin_f32 == nullptr && io.fin_u16 == nullptr &&
                          g.out_lane == 1 && g.out_outer % 8 == 0 && g.out_ax % 8 == 0 &&
                          reinterpret_cast<uintptr_t>(out) % 16 == 0 && io.out_lo % 8 == 0;
        if (!line || (io.widen != MXB_BF16 && io.widen != MXB_FP16)) return 0;
        return io.widen == MXB_BF16 ? 8 : 9;
    }
    if (mask != nullptr && bits.w == nullptr) return 0;
    const int fin = (io.fin_f32 != nullptr ? 1 : 0) + (io.fin_u16 != nullptr ? 2 : 0);
    if (fin != 0) {
        if (g.out_ax != 1 || g.out_lane % 8 != 0 || g.out_outer % 8 != 0) return 0;
        if (io.fin_f32 && reinterpret_cast<uintptr_t>(io.fin_f32) % 16 != 0) return 0;
        if (io.fin_u16 && reinterpret_cast<uintptr_t>(io.fin_u16) % 16 != 0) return 0;
        return 4 + fin;
    }
    if (reinterpret_cast<uintptr_t>(out) % 16 != 0 || io.out_lo % 8 != 0) return 0;
    if (mask == nullptr && g.out_lane == 1 && g.out_outer % 8 == 0 && g.out_ax % 8 == 0) return 1;
    if (g.out_ax == 1 && g.out_lane % 8 == 0 && g.out_outer % 8 == 0) return mask == nullptr ? 2 : 3;
    return 0;
}

template <int RT>
static hipError_t launch_split_rt(int form, const uint16_t *in, uint16_t *out, const uint8_t *mask,
                                  const MaskBits &bits, LineGeom g, const float *k, int R, SplitIO io,
                                  hipStream_t s) {
    const int64_t groups = ((g.n_lane + 63) / 64) * g.n_outer;
    const int64_t target = 256LL * 6;
    int64_t nseg = groups > 0 ? (target + groups - 1) / groups : 1;
    const int64_t max_seg = (g.n_ax + 63) / 64;
    if (nseg > max_seg) nseg = max_seg;
    if (nseg < 1) nseg = 1;
    int64_t seg = (g.n_ax + nseg - 1) / nseg;
    seg = ((seg + 31) / 32) * 32;
    g.seg = seg;
    g.nseg = (g.n_ax + seg - 1) / seg;
    const int64_t blocks = g.nseg * groups;
    if (blocks == 0) return hipSuccess;
    float scale = 1.0f;
    const TapsL<2 * RT + 1> t = split_taps_scaled<RT>(k, R, &scale);
    io.scale = scale;
    const dim3 grid((unsigned)blocks), block(256);
#ifndef MXB_SPLIT_DZ64
#define MXB_SPLIT_DZ64 2
#endif
    // the z-pass of radius bucket 64 prefetches two pairs, not three, so that
    // two blocks share a CU (76.5 against 84.7 KB of LDS): 10.86 against
    // 12.80 ms per launch on 2048x2048x1536, R=60
    constexpr int DZ = RT > 48 && MXB_SPLIT_DEPTH_Z > MXB_SPLIT_DZ64 ? MXB_SPLIT_DZ64 : MXB_SPLIT_DEPTH_Z;
    constexpr int DY = MXB_SPLIT_DEPTH_Y, DX = MXB_SPLIT_DEPTH_X;
    switch (form) {
    case 1: hipLaunchKernelGGL((k_line_split_dma<RT, DZ, false, true, 0>), grid, block, 0, s, in, out, mask, bits, g, t, io); break;
    case 2: hipLaunchKernelGGL((k_line_split_dma<RT, DY, false, false, 0>), grid, block, 0, s, in, out, mask, bits, g, t, io); break;
    case 3: hipLaunchKernelGGL((k_line_split_dma<RT, DX, true, false, 0>), grid, block, 0, s, in, out, mask, bits, g, t, io); break;
    case 5: hipLaunchKernelGGL((k_line_split_dma<RT, DX, true, false, 1>), grid, block, 0, s, in, out, mask, bits, g, t, io); break;
    case 6: hipLaunchKernelGGL((k_line_split_dma<RT, DX, true, false, 2>), grid, block, 0, s, in, out, mask, bits, g, t, io); break;
    case 7: hipLaunchKernelGGL((k_line_split_dma<RT, DX, true, false, 3>), grid, block, 0, s, in, out, mask, bits, g, t, io); break;
    case 8: hipLaunchKernelGGL((k_line_split_dma<RT, DZ, false, true, 0, 1>), grid, block, 0, s, in, out, mask, bits, g, t, io); break;
    case 9: hipLaunchKernelGGL((k_line_split_dma<RT, DZ, false, true, 0, 2>), grid, block, 0, s, in, out, mask, bits, g, t, io); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

bool split_pass_supported(const uint16_t *in, const uint16_t *out, const uint8_t *mask, const MaskBits &bits,
                          const LineGeom &g, int R, const SplitIO &io) {
    // the final-output forms always illuminate (they end a repetition)
    if (io.fin_f32 != nullptr || io.fin_u16 != nullptr) {
        if (mask == nullptr) return false;
    }
    return R <= kMaxRadiusLowp && split_form(in, out, mask, bits, g, io) != 0;
}

hipError_t launch_line_pass_split(const uint16_t *in, uint16_t *out, const uint8_t *mask, const MaskBits &bits,
                                  const LineGeom &g, const float *taps, int R, const SplitIO &io,
                                  hipStream_t s) {
    if (!split_pass_supported(in, out, mask, bits, g, R, io)) return hipErrorInvalidValue;
    const int form = split_form(in, out, mask, bits, g, io);
#define MXB_SPLIT_CASE(rt) \
    if (R <= rt) return launch_split_rt<rt>(form, in, out, mask, bits, g, taps, R, io, s);
    // the 16-bit forms' radius buckets
    MXB_SPLIT_CASE(4)
    MXB_SPLIT_CASE(8)
    MXB_SPLIT_CASE(12)
    MXB_SPLIT_CASE(16)
    MXB_SPLIT_CASE(24)
    MXB_SPLIT_CASE(32)
    MXB_SPLIT_CASE(40)
    MXB_SPLIT_CASE(48)
    MXB_SPLIT_CASE(64)
#undef MXB_SPLIT_CASE
    return hipErrorInvalidValue;
}

bool lowp_lds_supported(const uint16_t *in, const LineGeom &g, int R) {
    return R <= 64 && g.in_outer % 8 == 0 && g.in_ax % 8 == 0 &&
           reinterpret_cast<uintptr_t>(in) % 16 == 0;
}

hipError_t launch_line_pass_lowp_lds(const uint16_t *in, uint16_t *out, const uint8_t *mask,
                                     const MaskBits &bits, const LineGeom &g, const float *taps,
                                     int R, bool bf, hipStream_t s) {
#define MXB_LOWP_LDS_CASE(rt) \
    if (R <= rt) return launch_lowp_staged_rt<rt>(in, out, mask, bits, g, taps, R, bf, s);
    // the same radius buckets as the register-ring kernel: the bucket fixes
    // how taps group into 32-position chunks, so both forms agree bit for bit
    MXB_LOWP_LDS_CASE(4)
    MXB_LOWP_LDS_CASE(8)
    MXB_LOWP_LDS_CASE(12)
    MXB_LOWP_LDS_CASE(16)
    MXB_LOWP_LDS_CASE(24)
    MXB_LOWP_LDS_CASE(32)
    MXB_LOWP_LDS_CASE(40)
    MXB_LOWP_LDS_CASE(48)
    MXB_LOWP_LDS_CASE(64)
#undef MXB_LOWP_LDS_CASE
    return hipErrorInvalidValue;
}

}  // namespace mxb
