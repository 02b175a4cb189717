// Internal declarations shared by the HIP translation units of libmxb_diffusion.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>

namespace mxb {

// Status codes of the C ABI (mirrored in include/maxibone_diffusion.h).
enum Status : int { OK = 0, EINVAL_ = 1, EDEVICE = 2, ENOMEM_ = 3, EIO_ = 4 };

void set_error(const std::string &msg);

// Largest radius served by the register-window kernels; larger radii use the
// per-voxel kernel, up to kMaxTapsGeneric taps.
constexpr int kMaxRadiusFast = 64;
constexpr int kMaxTapsGeneric = 511;

// Geometry of one 1D pass.  A "line" is the set of voxels a 1D correlation
// runs along.  Lines are grouped so that 64 consecutive lines sit at
// consecutive addresses of the input (lane stride 1): every wave reads 256
// contiguous bytes per position.  `outer` indexes groups of lanes.
struct LineGeom {
    int64_t n_lane, n_outer, n_ax;   // lines per group, groups, positions per line
    int64_t in_outer, in_ax;         // input strides (lane stride is 1)
    int64_t out_lane, out_outer, out_ax;  // output strides (also the mask's)
    int64_t seg;                     // positions per thread segment (multiple of 16)
    int64_t nseg;                    // segments per line
    int32_t vec_out;                 // 1: out_ax == 1 and every output row is 16-byte aligned
    int32_t stage_out;               // 1: out_lane == 1, rows 16-byte aligned: stage tiles in LDS
};

// Host <-> device copy of `bytes`, ordered on `s` after its queued work and
// complete on return; copies of 128 MB and more between pageable host memory
// and the device go through pinned chunks filled or drained by host threads.
hipError_t staged_copy(void *dst, const void *src, size_t bytes, bool h2d, hipStream_t s);

// The mask packed to bits along x, words interleaved across y: bit (x mod 32)
// of word [(z * words_per_row(nx) + x / 32) * ny + y] is (mask[z][y][x] != 0),
// so one 32-position word of consecutive lines (y) is contiguous.  The
// illuminating pass's direct-load forms read it (1/8 byte per voxel instead of
// 1 byte); every other form reads the byte mask.
struct MaskBits {
    const uint32_t *w = nullptr;            // nullptr: not available
    int64_t lane = 0, outer = 0, word = 0;  // word strides: the pass's lines, outer index, 32 positions
};
inline int64_t mask_words_per_row(int64_t nx) { return (nx + 31) / 32; }
inline int64_t mask_bits_bytes(int64_t nz, int64_t ny, int64_t nx) {
    return nz * ny * mask_words_per_row(nx) * 4;
}
// Packs nz planes of ny rows of nx mask bytes (row stride mask_row, plane
// stride mask_plane) into the layout above.
hipError_t launch_pack_mask_bits(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx, int64_t mask_row,
                                 int64_t mask_plane, uint32_t *bits, hipStream_t s);

// Epilogue of the last pass of a step: illuminate from the mask, write the
// float state and/or its uint16 quantisation.
struct Epilogue {
    const uint8_t *mask;   // nullptr: plain pass (no illumination)
    float *out_f32;        // may be nullptr when only out_u16 is wanted
    uint16_t *out_u16;     // nullptr unless this is the final step
    MaskBits bits;         // the same mask as bits, when available
};

enum class Arith : int { FMA = 0, EXACT = 1, MFMA = 2 };

// MXB_PASS_KERNEL=ring selects the register-ring MFMA kernels even where the
// LDS-staged ones apply (both compute the same values bit for bit; the switch
// lets tests compare them).  Read at every launch.
// MXB_PASS_KERNEL=staged selects, among the LDS-staged kernels, the form that
// stages loads through VGPRs instead of direct global -> LDS loads.
inline bool force_ring_kernels() {
    const char *v = std::getenv("MXB_PASS_KERNEL");
    return v != nullptr && std::strcmp(v, "ring") == 0;
}
inline bool force_vgpr_staging() {
    const char *v = std::getenv("MXB_PASS_KERNEL");
    return v != nullptr && std::strcmp(v, "staged") == 0;
}

// Pass launchers (mxb_kernels_f32.hip).  `taps` are K = 2R+1 host floats.
hipError_t launch_line_pass_f32(const float *in, const LineGeom &g, const Epilogue &e,
                                const float *taps, int R, Arith arith, hipStream_t s);
hipError_t launch_init_state(const uint8_t *mask, float *u, int64_t n, hipStream_t s);
hipError_t launch_quantise(const float *u, uint16_t *q, float *f, int64_t n, hipStream_t s);
// m = (u == 1.0f): the reference on-disk path's per-chunk mask rule.
hipError_t launch_mask_from_state(const float *u, uint8_t *m, int64_t n, hipStream_t s);

// LDS-staged FP32 MFMA pass (mxb_kernels_f32_lds.hip): 16-byte loads of whole
// rows; needs 16-byte-aligned input rows.
bool lds_pass_supported(const float *in, const LineGeom &g, int R);
hipError_t launch_line_pass_f32_lds(const float *in, const LineGeom &g, const Epilogue &e,
                                    const float *taps, int R, hipStream_t s);

// Reduced-precision passes (mxb_kernels_lowp.hip): 16-bit state in and out,
// `bf` selects BF16 (else FP16).  `mask` non-null illuminates.
hipError_t launch_line_pass_lowp(const uint16_t *in, uint16_t *out, const uint8_t *mask,
                                 const LineGeom &g, const float *taps, int R, bool bf,
                                 hipStream_t s, MaskBits bits = MaskBits{});
hipError_t launch_init_lowp(const uint8_t *mask, uint16_t *u, int64_t n, bool bf, hipStream_t s);
hipError_t launch_lowp_to_f32(const uint16_t *u, float *f, int64_t n, bool bf, hipStream_t s);
constexpr int kMaxRadiusLowp = 64;

// Extra operands of the split (FP32-class) form.  Its state travels as two
// FP16 planes, u = hi + lo with hi = f16(u) and lo = f16(u - hi): the low
// plane sits `in_lo` / `out_lo` elements after the high plane's start.  Each
// chunk runs three MFMAs, hi * k_hi, hi * k_lo and lo * k_hi (the dropped
// lo * k_lo is below 2^-22 of the product), on taps pre-scaled by a power of
// two so that their low parts stay normal FP16 numbers; the accumulator times
// `scale` (the inverse power of two, exact) is the FP32 value.  fin_f32 /
// fin_u16: the final pass writes FP32 and/or uint16 outputs at the pass's
// output strides instead of the two planes.
struct SplitIO {
    int64_t in_lo = 0, out_lo = 0;
    float scale = 1.0f;
    float *fin_f32 = nullptr;
    uint16_t *fin_u16 = nullptr;
    // z-pass only: the input is one 16-bit state plane (MXB_BF16 or
    // MXB_FP16; 0 = hi + lo planes), taken as the hi plane with lo 0 (BF16
    // values widened to FP16), so a 16-bit state enters the split form
    int widen = 0;
};

// Split form (mxb_kernels_lowp_lds.hip, mxb_lowp_dma_body.h).  A pass reads
// the FP16 hi + lo planes at `in` (lo io.in_lo elements further) and writes
// planes at `out` (lo io.out_lo further), or, when io.fin_f32 / io.fin_u16
// are set, the final outputs; `mask` non-null illuminates.  io.scale is set
// by the launcher.
bool split_pass_supported(const uint16_t *in, const uint16_t *out, const uint8_t *mask, const MaskBits &bits,
                          const LineGeom &g, int R, const SplitIO &io);
hipError_t launch_line_pass_split(const uint16_t *in, uint16_t *out, const uint8_t *mask, const MaskBits &bits,
                                  const LineGeom &g, const float *taps, int R, const SplitIO &io,
                                  hipStream_t s);
// Fused x -> y pass of the 16-bit forms (mxb_fused.hip): the z-pass output
// `in` (one 16-bit plane, or the split form's two) -> `out` with
// illumination from the mask bits (and, for the split form with io.fin_*
// set, the final outputs instead of planes).  precision: MXB_BF16, MXB_FP16 or
// MXB_FP32_SPLIT.
bool xy_fused_supported(const uint16_t *in, const uint16_t *out, const MaskBits &bits, int64_t nz, int64_t ny,
                        int64_t nx, int R, int precision, const SplitIO &io);
hipError_t launch_xy_fused(const uint16_t *in, uint16_t *out, const MaskBits &bits, int64_t nz, int64_t ny,
                           int64_t nx, const float *taps, int R, int precision, const SplitIO &io, hipStream_t s);
// MXB_XY_FUSED=0 runs the three-pass form instead (read at every call; the
// pass-kernel selection MXB_PASS_KERNEL does too).
inline bool xy_fusion_enabled() {
    const char *v = std::getenv("MXB_XY_FUSED");
    if (v != nullptr && std::strcmp(v, "0") == 0) return false;
    return !force_ring_kernels() && !force_vgpr_staging();
}

// Split-form state initialisation: planes from the mask (hi = 1 or 0, lo = 0)
// or from an FP32 state (hi = f16(u), lo = f16(u - hi)); n voxels, the low
// plane lo_off elements after the high one.
hipError_t launch_init_split(const uint8_t *mask, uint16_t *planes, int64_t lo_off, int64_t n, hipStream_t s);
hipError_t launch_split_state(const float *u, uint16_t *planes, int64_t lo_off, int64_t n, hipStream_t s);

// mxb_device_diffusion with the pass structure chosen: fused = false keeps
// the three-pass form of the 16-bit and split precisions (the structure of
// the batched persistent passes, so per-volume runs compare bit for bit).
int device_diffusion(const uint8_t *dev_mask, int64_t nz, int64_t ny, int64_t nx, const float *kernel,
                     int64_t kernel_size, int64_t repetitions, int precision, void *dev_state0,
                     void *dev_state1, uint16_t *dev_out_u16, float *dev_out_f32, void *stream, bool fused);

// Plan the three passes of a step over a (nz, ny, nx) volume.
// z: A -> A, y: A -> T (T is (z, x, y) order), x: T -> A.
LineGeom geom_z(int64_t nz, int64_t ny, int64_t nx);
LineGeom geom_y(int64_t nz, int64_t ny, int64_t nx);
LineGeom geom_x(int64_t nz, int64_t ny, int64_t nx);

}  // namespace mxb
