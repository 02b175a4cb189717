// FP32 line pass, LDS-staged forms: single-volume kernels and launcher (the
// per-block bodies live in mxb_lds_body.h and mxb_f32_dma_body.h).
#include "mxb_f32_dma_body.h"

namespace mxb {

template <int RT>
__global__ __launch_bounds__(256) void k_line_mfma_lds(const float *__restrict__ in, LineGeom g,
                                                       Epilogue e, TapsF<2 * RT + 1> taps) {
    __shared__ float smem[LdsSmem<RT>::TOTAL];
    lds_fill_taps<RT>(smem, taps.k);
    int64_t bid = blockIdx.x;
    const int64_t seg_id = bid % g.nseg;
    bid /= g.nseg;
    const int64_t nlb = (g.n_lane + 63) >> 6;
    const int64_t lb = bid % nlb;
    const int64_t outer = bid / nlb;
    lds_line_body<RT>(in, g, e, outer, lb, seg_id, smem);
}

template <int RT, int D, bool MASK, bool OUTF, bool OUTU, bool ZSTAGE>
__global__ __launch_bounds__(256) void k_line_f32_dma(const float *__restrict__ in, LineGeom g,
                                                      Epilogue e, TapsF<2 * RT + 1> taps) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[F32DmaForm<RT, D, MASK, ZSTAGE>::BYTES];
    int64_t bid = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t seg_id = bid % g.nseg;
    bid /= g.nseg;
    const int64_t nlb = (g.n_lane + 63) >> 6;
    const int64_t lb = bid % nlb;
    const int64_t outer = bid / nlb;
    f32_dma_pass<RT, D, MASK, OUTF, OUTU, ZSTAGE>(in, g, e, outer, lb, seg_id, taps.k, smem);
}

// Which direct-load epilogue serves (g, e): 0 none (use the VGPR-staged
// form), else a code for the template below.
static int f32_dma_form(const LineGeom &g, const Epilogue &e) {
    if (g.n_lane % 4 != 0) return 0;
    const bool f = e.out_f32 != nullptr, u = e.out_u16 != nullptr, m = e.mask != nullptr;
    if (f && reinterpret_cast<uintptr_t>(e.out_f32) % 16 != 0) return 0;
    if (u && reinterpret_cast<uintptr_t>(e.out_u16) % 16 != 0) return 0;
    if (g.out_lane == 1 && !m && !u && f && g.out_outer % 4 == 0 && g.out_ax % 4 == 0) return 1;
    if (g.out_ax != 1 || g.out_lane % 8 != 0 || g.out_outer % 8 != 0) return 0;
    if (m && e.bits.w == nullptr) return 0;  // the direct-load illuminating pass reads mask bits
    return 2 + (m ? 4 : 0) + (f ? 1 : 0) + (u ? 2 : 0);
}

template <int RT>
static hipError_t launch_f32_dma_rt(int form, const float *in, LineGeom g, const Epilogue &e,
                                    const TapsF<2 * RT + 1> &t, hipStream_t s) {
    constexpr int D = MXB_F32_DMA_DEPTH;
    const dim3 grid((unsigned)(g.nseg * ((g.n_lane + 63) / 64) * g.n_outer)), block(256);
    switch (form) {
    case 1: hipLaunchKernelGGL((k_line_f32_dma<RT, D, false, true, false, true>), grid, block, 0, s, in, g, e, t); break;
    case 3: hipLaunchKernelGGL((k_line_f32_dma<RT, D, false, true, false, false>), grid, block, 0, s, in, g, e, t); break;
    case 4: hipLaunchKernelGGL((k_line_f32_dma<RT, D, false, false, true, false>), grid, block, 0, s, in, g, e, t); break;
    case 5: hipLaunchKernelGGL((k_line_f32_dma<RT, D, false, true, true, false>), grid, block, 0, s, in, g, e, t); break;
    case 7: hipLaunchKernelGGL((k_line_f32_dma<RT, D, true, true, false, false>), grid, block, 0, s, in, g, e, t); break;
    case 8: hipLaunchKernelGGL((k_line_f32_dma<RT, D, true, false, true, false>), grid, block, 0, s, in, g, e, t); break;
    case 9: hipLaunchKernelGGL((k_line_f32_dma<RT, D, true, true, true, false>), grid, block, 0, s, in, g, e, t); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int RT>
static TapsF<2 * RT + 1> centred_taps_l(const float *k, int R) {
    TapsF<2 * RT + 1> t;
    for (int i = 0; i < 2 * RT + 1; ++i) {
        const int r = i - RT;
        t.k[i] = (r >= -R && r <= R) ? k[R + r] : 0.0f;
    }
    return t;
}

template <int RT>
static hipError_t launch_lds_rt(const float *in, LineGeom g, const Epilogue &e, const float *k,
                                int R, hipStream_t s) {
    // one block per (segment, 64 lines, outer); segments keep >= ~16 waves/CU
    const int64_t groups = ((g.n_lane + 63) / 64) * g.n_outer;
    const int64_t target = 256LL * 6;
    int64_t nseg = groups > 0 ? (target + groups - 1) / groups : 1;
    const int64_t max_seg = (g.n_ax + 63) / 64;
    if (nseg > max_seg) nseg = max_seg;
    if (nseg < 1) nseg = 1;
    int64_t seg = (g.n_ax + nseg - 1) / nseg;
    const int64_t q = e.mask != nullptr ? 32 : 16;  // the illuminating pass: whole mask-bit words
    seg = ((seg + q - 1) / q) * q;
    g.seg = seg;
    g.nseg = (g.n_ax + seg - 1) / seg;
    g.vec_out = (g.out_ax == 1 && g.out_lane % 4 == 0 && g.out_outer % 4 == 0 &&
                 (e.out_f32 == nullptr || reinterpret_cast<uintptr_t>(e.out_f32) % 16 == 0) &&
                 (e.out_u16 == nullptr || reinterpret_cast<uintptr_t>(e.out_u16) % 8 == 0) &&
                 (e.mask == nullptr || reinterpret_cast<uintptr_t>(e.mask) % 4 == 0)) ? 1 : 0;
    g.stage_out = (g.out_lane == 1 && g.out_outer % 4 == 0 && g.out_ax % 4 == 0 &&
                   e.mask == nullptr && e.out_u16 == nullptr && e.out_f32 != nullptr &&
                   reinterpret_cast<uintptr_t>(e.out_f32) % 16 == 0) ? 1 : 0;
    const int64_t blocks = g.nseg * groups;
    if (blocks == 0) return hipSuccess;
    const TapsF<2 * RT + 1> t = centred_taps_l<RT>(k, R);
    const int form = force_vgpr_staging() ? 0 : f32_dma_form(g, e);
    if (form != 0) return launch_f32_dma_rt<RT>(form, in, g, e, t, s);
    hipLaunchKernelGGL((k_line_mfma_lds<RT>), dim3((unsigned)blocks), dim3(256), 0, s, in, g, e, t);
    return hipGetLastError();
}

// Usable when every input row is 16-byte aligned (in_outer, in_ax multiples of
// 4 floats) and the radius is compiled.
bool lds_pass_supported(const float *in, const LineGeom &g, int R) {
    return R <= 64 && g.in_outer % 4 == 0 && g.in_ax % 4 == 0 &&
           reinterpret_cast<uintptr_t>(in) % 16 == 0;
}

hipError_t launch_line_pass_f32_lds(const float *in, const LineGeom &g, const Epilogue &e,
                                    const float *taps, int R, hipStream_t s) {
#define MXB_LDS_CASE(rt) \
    if (R <= rt) return launch_lds_rt<rt>(in, g, e, taps, R, s);
    MXB_LDS_CASE(4)
    MXB_LDS_CASE(8)
    MXB_LDS_CASE(12)
    MXB_LDS_CASE(16)
    MXB_LDS_CASE(20)
    MXB_LDS_CASE(24)
    MXB_LDS_CASE(32)
    MXB_LDS_CASE(40)
    MXB_LDS_CASE(48)
    MXB_LDS_CASE(56)
    MXB_LDS_CASE(64)
#undef MXB_LDS_CASE
    return hipErrorInvalidValue;
}

}  // namespace mxb
