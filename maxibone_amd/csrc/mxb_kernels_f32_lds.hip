// FP32 line pass, LDS-staged form: single-volume kernel and launcher (the
// per-block body lives in mxb_lds_body.h).
#include "mxb_lds_body.h"

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
    seg = ((seg + 15) / 16) * 16;
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
