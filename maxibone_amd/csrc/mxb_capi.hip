// C ABI of libmxb_diffusion: argument checks, device buffers, the repetition
// driver, the z-chunked and file-backed forms, and kernel timing.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "maxibone_diffusion.h"
#include "mxb_internal.h"

namespace mxb {

static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }

#define MXB_HIP(expr)                                                                   \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess) {                                                         \
            set_error(std::string(#expr) + ": " + hipGetErrorString(_e));               \
            return _e == hipErrorOutOfMemory ? ENOMEM_ : EDEVICE;                       \
        }                                                                               \
    } while (0)

// ---------------------------------------------------------------- timing
struct TimingState {
    std::mutex mu;
    bool enabled = false;
    struct Rec {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double total_ms[5] = {0, 0, 0, 0, 0};
    int64_t launches[5] = {0, 0, 0, 0, 0};
};
static TimingState g_timing;

static hipEvent_t take_event() {
    if (!g_timing.pool.empty()) {
        hipEvent_t e = g_timing.pool.back();
        g_timing.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Set while a call must not create events (the no-allocation entry point).
static thread_local bool g_timing_off = false;

// Runs `fn` (a launch on stream s), bracketing it with events when timing is on.
template <typename F>
static hipError_t timed(int cls, hipStream_t s, F &&fn) {
    if (!g_timing.enabled || g_timing_off) return fn();
    std::lock_guard<std::mutex> lk(g_timing.mu);
    hipEvent_t a = take_event(), b = take_event();
    if (a == nullptr || b == nullptr) return fn();
    hipEventRecord(a, s);
    hipError_t e = fn();
    hipEventRecord(b, s);
    g_timing.recs.push_back({cls, a, b});
    return e;
}

static void drain_timing_locked() {
    for (auto &r : g_timing.recs) {
        float ms = 0.0f;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            g_timing.total_ms[r.cls] += ms;
            g_timing.launches[r.cls] += 1;
        }
        g_timing.pool.push_back(r.a);
        g_timing.pool.push_back(r.b);
    }
    g_timing.recs.clear();
}

// ---------------------------------------------------------------- checks
static int check_common(int64_t nz, int64_t ny, int64_t nx, const float *kernel, int64_t K,
                        int64_t reps, int precision) {
    if (nz < 0 || ny < 0 || nx < 0) {
        set_error("volume dimensions must be non-negative");
        return EINVAL_;
    }
    if (kernel == nullptr || K < 1 || (K % 2) == 0) {
        set_error("kernel must hold an odd number of taps (got " + std::to_string(K) + ")");
        return EINVAL_;
    }
    if (K > kMaxTapsGeneric) {
        set_error("kernel_size " + std::to_string(K) + " exceeds the supported maximum of " +
                  std::to_string(kMaxTapsGeneric));
        return EINVAL_;
    }
    if (reps < 0) {
        set_error("repetitions must be non-negative");
        return EINVAL_;
    }
    if (precision < MXB_FP32 || precision > MXB_FP32_SPLIT) {
        set_error("unknown precision " + std::to_string(precision));
        return EINVAL_;
    }
    if (nz > 0 && ny > 0 && nx > 0 && (nz > (int64_t(1) << 40) / ny / nx)) {
        set_error("volume too large");
        return EINVAL_;
    }
    return OK;
}

static Arith arith_of(int precision) {
    switch (precision) {
        case MXB_FP32_FMA: return Arith::FMA;
        case MXB_FP32_EXACT: return Arith::EXACT;
        default: return Arith::MFMA;
    }
}

static bool is_lowp(int precision) { return precision == MXB_BF16 || precision == MXB_FP16; }

// Every form needs two FP32-sized state buffers: the 16-bit forms use the
// first half of each for their repetitions and the whole of both for the
// final repetition (FP32 state, or the split planes and the FP32 result);
// the split form holds its FP16 high plane at the
// start and its low plane split_lo_off(n) elements further (4 KB past its
// end).  The first
// buffer also carries the mask packed to bits after the state (at byte
// state_f32_bytes).
// The low plane starts kSplitLoGap elements past the high plane's end
// rounded up to 128.  Without the gap the two planes of a power-of-two volume
// (1024^3: 2 GiB of FP16 each) sit a power of two apart, a lane's loads from
// both alias in HBM, and the split z-pass ran 7 % slower (1.92 against 1.78
// ms at 1024^3; a 1 MiB gap recovered half of that).
constexpr int64_t kSplitLoGap = 2048;
static int64_t split_lo_off(int64_t n) { return (n + 127) / 128 * 128 + kSplitLoGap; }
static int64_t state_f32_bytes(int64_t n) { return split_lo_off(n) * 4; }

int64_t state_bytes(int precision, int64_t nz, int64_t ny, int64_t nx) {
    (void)precision;
    return state_f32_bytes(nz * ny * nx) + mask_bits_bytes(nz, ny, nx);
}

// Packs the (nz, ny, nx) device mask to bits at `w` and describes them for
// the x-pass (lines y, outer index z).
static hipError_t pack_bits(const uint8_t *d_mask, int64_t nz, int64_t ny, int64_t nx, uint32_t *w,
                            MaskBits &bits, hipStream_t s) {
    const int64_t wpr = mask_words_per_row(nx);
    bits = MaskBits{w, 1, wpr * ny, ny};
    return launch_pack_mask_bits(d_mask, nz, ny, nx, nx, ny * nx, w, s);
}

// ---------------------------------------------------------------- driver
// Runs `reps` repetitions on device buffers.  If init_from_mask, state0 is
// first set to (mask > 0); otherwise state0 already holds the FP32 state.
// FP32 step: z: s0 -> s1, y: s1 -> s0 ((z, x, y) order), x: s0 -> s1 with
// illumination; then the buffers swap roles.  The final x-pass writes
// out_f32 / out_u16 (when non-null); when out_f32 is null the final FP32
// state is left in the buffer returned through *final_state (may be null).
static int run_steps_f32(const uint8_t *d_mask, const MaskBits &bits, int64_t nz, int64_t ny,
                         int64_t nx, const float *k, int64_t K, int64_t reps, Arith arith, float *s0,
                         float *s1, bool init_from_mask, uint16_t *out_u16, float *out_f32,
                         float **final_state, hipStream_t s) {
    const int64_t n = nz * ny * nx;
    const int R = (int)(K / 2);
    if (n == 0) return OK;
    if (init_from_mask) MXB_HIP(timed(MXB_K_OTHER, s, [&] { return launch_init_state(d_mask, s0, n, s); }));
    float *cur = s0, *tmp = s1;
    if (reps == 0) {
        if (out_u16 != nullptr || out_f32 != nullptr)
            MXB_HIP(timed(MXB_K_OTHER, s, [&] { return launch_quantise(cur, out_u16, out_f32, n, s); }));
        if (final_state) *final_state = cur;
        return OK;
    }
    const LineGeom gz = geom_z(nz, ny, nx), gy = geom_y(nz, ny, nx), gx = geom_x(nz, ny, nx);
    for (int64_t t = 0; t < reps; ++t) {
        const bool last = t == reps - 1;
        const Epilogue ez{nullptr, tmp, nullptr, MaskBits{}};
        MXB_HIP(timed(MXB_K_ZPASS, s, [&] { return launch_line_pass_f32(cur, gz, ez, k, R, arith, s); }));
        const Epilogue ey{nullptr, cur, nullptr, MaskBits{}};
        MXB_HIP(timed(MXB_K_YPASS, s, [&] { return launch_line_pass_f32(tmp, gy, ey, k, R, arith, s); }));
        // the last x-pass writes the float state only where someone reads it
        const bool state_unread = last && out_f32 == nullptr && out_u16 != nullptr && final_state == nullptr;
        float *dst = (last && out_f32 != nullptr) ? out_f32 : (state_unread ? nullptr : tmp);
        const Epilogue ex{d_mask, dst, last ? out_u16 : nullptr, bits};
        MXB_HIP(timed(MXB_K_XPASS, s, [&] { return launch_line_pass_f32(cur, gx, ex, k, R, arith, s); }));
        if (last && out_f32 != nullptr) {
            if (final_state) *final_state = out_f32;
        } else {
            std::swap(cur, tmp);
            if (last && final_state) *final_state = cur;
        }
    }
    return OK;
}

// 16-bit repetitions: n16 repetitions keep BF16/FP16 state (z: s0 -> s1,
// y: s1 -> s0, x: s0 -> s1 + illumination, swap), starting from
// (mask > 0) when init_from_mask, else from the 16-bit state already in s0.
// With final_f32 one more repetition then runs at FP32 accuracy, writing
// out_u16 / out_f32, so the result carries no 16-bit rounding of its own
// repetition: where the fused pass serves the volume, in the split form (its
// z-pass reads the 16-bit state as the FP16 high plane, the fused x -> y
// pass writes the outputs); otherwise, and for a lone repetition from the
// mask (as in-memory runs of one repetition do), the state is widened to
// FP32 for one FP32 (MFMA) repetition.  The choice depends only on the
// volume and the 16-bit state, so chunked runs, which visit the final
// repetition with n16 = 0, match in-memory runs bit for bit.  Without
// final_f32 the 16-bit state is left where *final16 points.
static int run_lowp_core(const uint8_t *d_mask, const MaskBits &bits, int64_t nz, int64_t ny,
                         int64_t nx, const float *k, int64_t K, int64_t n16, bool final_f32, bool bf,
                         bool init_from_mask, void *s0, void *s1, uint16_t *out_u16, float *out_f32,
                         float **final_state, uint16_t **final16, bool fused, hipStream_t s) {
    const int64_t n = nz * ny * nx;
    const int R = (int)(K / 2);
    if (n == 0) return OK;
    uint16_t *cur = (uint16_t *)s0, *tmp = (uint16_t *)s1;
    if (init_from_mask) MXB_HIP(timed(MXB_K_OTHER, s, [&] { return launch_init_lowp(d_mask, cur, n, bf, s); }));
    const LineGeom gz = geom_z(nz, ny, nx), gy = geom_y(nz, ny, nx), gx = geom_x(nz, ny, nx);
    const int prec = bf ? MXB_BF16 : MXB_FP16;
    fused = fused && xy_fusion_enabled() && xy_fused_supported(tmp, cur, bits, nz, ny, nx, R, prec, SplitIO{});
    for (int64_t t = 0; t < n16; ++t) {
        MXB_HIP(timed(MXB_K_ZPASS, s, [&] { return launch_line_pass_lowp(cur, tmp, nullptr, gz, k, R, bf, s); }));
        if (fused) {
            // x then y in one pass, back into cur, illuminating
            MXB_HIP(timed(MXB_K_XYPASS, s, [&] {
                return launch_xy_fused(tmp, cur, bits, nz, ny, nx, k, R, prec, SplitIO{}, s);
            }));
            continue;
        }
        MXB_HIP(timed(MXB_K_YPASS, s, [&] { return launch_line_pass_lowp(tmp, cur, nullptr, gy, k, R, bf, s); }));
        MXB_HIP(timed(MXB_K_XPASS, s, [&] { return launch_line_pass_lowp(cur, tmp, d_mask, gx, k, R, bf, s, bits); }));
        std::swap(cur, tmp);
    }
    if (!final_f32) {
        if (final16) *final16 = cur;
        return OK;
    }
    if (fused && !(init_from_mask && n16 == 0)) {
        // split planes into the other buffer; the 16-bit state's buffer then
        // takes the FP32 state when one is wanted
        const int64_t lo = split_lo_off(n);
        SplitIO iz;
        iz.out_lo = lo;
        iz.widen = prec;
        SplitIO ex;
        ex.in_lo = ex.out_lo = lo;
        ex.fin_u16 = out_u16;
        ex.fin_f32 = out_f32 != nullptr ? out_f32 : (final_state != nullptr || out_u16 == nullptr ? (float *)cur : nullptr);
        if (split_pass_supported(cur, tmp, nullptr, bits, gz, R, iz) &&
            xy_fused_supported(tmp, cur, bits, nz, ny, nx, R, MXB_FP32_SPLIT, ex)) {
            MXB_HIP(timed(MXB_K_ZPASS, s, [&] { return launch_line_pass_split(cur, tmp, nullptr, bits, gz, k, R, iz, s); }));
            MXB_HIP(timed(MXB_K_XYPASS, s, [&] {
                return launch_xy_fused(tmp, cur, bits, nz, ny, nx, k, R, MXB_FP32_SPLIT, ex, s);
            }));
            if (final_state) *final_state = ex.fin_f32;
            return OK;
        }
    }
    // widen into the other buffer, then one FP32 repetition
    float *f0 = (float *)(cur == (uint16_t *)s0 ? s1 : s0);
    float *f1 = (float *)(cur == (uint16_t *)s0 ? s0 : s1);
    MXB_HIP(timed(MXB_K_OTHER, s, [&] { return launch_lowp_to_f32(cur, f0, n, bf, s); }));
    return run_steps_f32(d_mask, bits, nz, ny, nx, k, K, 1, Arith::MFMA, f0, f1, false, out_u16,
                         out_f32, final_state, s);
}

// 16-bit form of a whole diffusion: repetitions 1 .. reps-1 in BF16/FP16,
// the last in FP32.
static int run_steps_lowp(const uint8_t *d_mask, const MaskBits &bits, int64_t nz, int64_t ny,
                          int64_t nx, const float *k, int64_t K, int64_t reps, bool bf, void *s0,
                          void *s1, uint16_t *out_u16, float *out_f32, bool fused, hipStream_t s) {
    if (reps <= 1)
        return run_steps_f32(d_mask, bits, nz, ny, nx, k, K, reps, Arith::MFMA, (float *)s0,
                             (float *)s1, true, out_u16, out_f32, nullptr, s);
    return run_lowp_core(d_mask, bits, nz, ny, nx, k, K, reps - 1, true, bf, true, s0, s1, out_u16,
                         out_f32, nullptr, nullptr, fused, s);
}

// How the split form serves a volume: 0 not at all (it runs as FP32), 1 with
// the three-pass form, 2 with the z-pass and the fused x -> y pass.
static int split_serves(int64_t nz, int64_t ny, int64_t nx, int R, const void *s0, const void *s1,
                        const MaskBits &bits, uint16_t *out_u16, float *out_f32, bool fused) {
    if (R > kMaxRadiusLowp || nz * ny * nx == 0) return 0;
    const int64_t lo = split_lo_off(nz * ny * nx);
    SplitIO io;
    io.in_lo = io.out_lo = lo;
    MaskBits b = bits;
    if (b.w == nullptr) b.w = reinterpret_cast<const uint32_t *>(s0);  // the passes will read bits
    const uint16_t *a = (const uint16_t *)s0, *c = (const uint16_t *)s1;
    const uint8_t *m = reinterpret_cast<const uint8_t *>(s0);
    SplitIO fin = io;  // the last pass writes the outputs, or the FP32 state into a state buffer
    fin.fin_u16 = out_u16;
    fin.fin_f32 = out_f32 != nullptr ? out_f32 : reinterpret_cast<float *>(const_cast<void *>(s0));
    if (!split_pass_supported(a, c, nullptr, b, geom_z(nz, ny, nx), R, io)) return 0;
    if (fused && xy_fusion_enabled() && xy_fused_supported(c, a, b, nz, ny, nx, R, MXB_FP32_SPLIT, io) &&
        xy_fused_supported(c, a, b, nz, ny, nx, R, MXB_FP32_SPLIT, fin))
        return 2;
    return split_pass_supported(c, a, nullptr, b, geom_y(nz, ny, nx), R, io) &&
                   split_pass_supported(a, c, m, b, geom_x(nz, ny, nx), R, io) &&
                   split_pass_supported(a, c, m, b, geom_x(nz, ny, nx), R, fin)
               ? 1
               : 0;
}

// Split form: reps repetitions on FP16 hi + lo planes, starting from the mask
// (from_f32 false) or from the FP32 state in s0 (split into s1 first).  Per
// repetition the z-pass writes the other buffer; then either the fused
// x -> y pass writes back (form 2) or the y-pass (in (z, x, y) order) and the
// x-pass do, swapping the buffers (form 1).  The last pass writes out_u16 /
// out_f32; with out_f32 null and final_state non-null it writes the FP32
// state into the buffer its input is not in and reports it.
static int run_steps_split(const uint8_t *d_mask, const MaskBits &bits, int64_t nz, int64_t ny, int64_t nx,
                           const float *k, int64_t K, int64_t reps, void *s0, void *s1, bool from_f32,
                           uint16_t *out_u16, float *out_f32, float **final_state, int form, hipStream_t s) {
    const int64_t n = nz * ny * nx;
    const int R = (int)(K / 2);
    if (n == 0) return OK;
    if (reps == 0)
        return run_steps_f32(d_mask, bits, nz, ny, nx, k, K, 0, Arith::MFMA, (float *)s0, (float *)s1,
                             !from_f32, out_u16, out_f32, final_state, s);
    const int64_t lo = split_lo_off(n);
    uint16_t *cur = (uint16_t *)s0, *tmp = (uint16_t *)s1;
    const LineGeom gz = geom_z(nz, ny, nx), gy = geom_y(nz, ny, nx), gx = geom_x(nz, ny, nx);
    SplitIO io;
    io.in_lo = io.out_lo = lo;
    // From the mask, the 0/1 state is exact in FP16: the first z-pass reads
    // it as the high plane alone (its low plane would be zeros).
    SplitIO iz0 = io;
    iz0.in_lo = 0;
    iz0.widen = MXB_FP16;
    const bool widen0 = !from_f32 && split_pass_supported(cur, tmp, nullptr, bits, gz, R, iz0);
    if (from_f32) {
        MXB_HIP(timed(MXB_K_OTHER, s, [&] { return launch_split_state((const float *)s0, tmp, lo, n, s); }));
        std::swap(cur, tmp);
    } else if (widen0) {
        MXB_HIP(timed(MXB_K_OTHER, s, [&] { return launch_init_lowp(d_mask, cur, n, false, s); }));
    } else {
        MXB_HIP(timed(MXB_K_OTHER, s, [&] { return launch_init_split(d_mask, cur, lo, n, s); }));
    }
    for (int64_t t = 0; t < reps; ++t) {
        const bool last = t == reps - 1;
        const SplitIO &iz = t == 0 && widen0 ? iz0 : io;
        MXB_HIP(timed(MXB_K_ZPASS, s, [&] { return launch_line_pass_split(cur, tmp, nullptr, bits, gz, k, R, iz, s); }));
        // the last pass's input is tmp (fused) or cur (three-pass): its other
        // buffer takes the FP32 state when one is wanted
        uint16_t *const free_buf = form == 2 ? cur : tmp;
        SplitIO ex = io;
        if (last) {
            ex.fin_u16 = out_u16;
            ex.fin_f32 = out_f32 != nullptr ? out_f32 : (final_state != nullptr ? (float *)free_buf : nullptr);
            if (ex.fin_f32 == nullptr && ex.fin_u16 == nullptr) ex.fin_f32 = (float *)free_buf;
            if (final_state) *final_state = ex.fin_f32;
        }
        if (form == 2) {
            MXB_HIP(timed(MXB_K_XYPASS, s, [&] {
                return launch_xy_fused(tmp, cur, bits, nz, ny, nx, k, R, MXB_FP32_SPLIT, ex, s);
            }));
            continue;
        }
        MXB_HIP(timed(MXB_K_YPASS, s, [&] { return launch_line_pass_split(tmp, cur, nullptr, bits, gy, k, R, io, s); }));
        MXB_HIP(timed(MXB_K_XPASS, s, [&] { return launch_line_pass_split(cur, tmp, d_mask, bits, gx, k, R, ex, s); }));
        std::swap(cur, tmp);
    }
    return OK;
}

static int run_steps(const uint8_t *d_mask, int64_t nz, int64_t ny, int64_t nx, const float *k,
                     int64_t K, int64_t reps, int precision, void *s0, void *s1,
                     uint16_t *out_u16, float *out_f32, hipStream_t s, bool fused = true) {
    MaskBits bits;
    if (reps > 0 && nz * ny * nx > 0) {
        uint32_t *w = reinterpret_cast<uint32_t *>(static_cast<char *>(s0) + state_f32_bytes(nz * ny * nx));
        MXB_HIP(timed(MXB_K_OTHER, s, [&] { return pack_bits(d_mask, nz, ny, nx, w, bits, s); }));
    }
    if (is_lowp(precision)) {
        if (K / 2 > kMaxRadiusLowp) {
            set_error("BF16/FP16 forms support kernel radius <= " + std::to_string(kMaxRadiusLowp));
            return EINVAL_;
        }
        return run_steps_lowp(d_mask, bits, nz, ny, nx, k, K, reps, precision == MXB_BF16, s0, s1,
                              out_u16, out_f32, fused, s);
    }
    if (precision == MXB_FP32_SPLIT) {
        const int form = split_serves(nz, ny, nx, (int)(K / 2), s0, s1, bits, out_u16, out_f32, fused);
        if (form != 0)
            return run_steps_split(d_mask, bits, nz, ny, nx, k, K, reps, s0, s1, false, out_u16, out_f32,
                                   nullptr, form, s);
    }
    return run_steps_f32(d_mask, bits, nz, ny, nx, k, K, reps, arith_of(precision), (float *)s0,
                         (float *)s1, true, out_u16, out_f32, nullptr, s);
}

// RAII device buffer.
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes > 0 ? bytes : 1); }
};

struct Stream {
    hipStream_t s = nullptr;
    ~Stream() {
        if (s) (void)hipStreamDestroy(s);
    }
};

static void log_v(int verbose, int level, const std::string &msg) {
    if (verbose >= level) {
        std::fprintf(stdout, "%s\n", msg.c_str());
        std::fflush(stdout);
    }
}

// Host-resident chunked driver.  hs0 / hs1: host state buffers of n elements
// each (float for the FP32 forms, 16-bit for BF16 / FP16; pinned, pageable or
// file-backed; unused, and may be null, when reps <= T).  The first block
// starts from u0 = (mask > 0) built on the device; the last writes only the
// outputs.  The device holds one chunk: layers [lo, hi) = [z0 - H, z0 + cz + H)
// clipped to the volume, H = T*R.  Clipping keeps the true zero boundary at
// the volume ends; at interior chunk edges the fake boundary corrupts at most
// R layers per repetition, so after T repetitions the interior [z0, z0 + cz)
// is exact.  The 16-bit forms round lo down to a multiple of 32: a 16-bit
// pass groups a voxel's taps by its place in a 16-position tile, and with
// that alignment every layer keeps its place, so 16-bit chunked results equal
// in-memory ones bit for bit (FP32 passes do not depend on the tile place).
// Each block of a 16-bit diffusion runs its repetitions in 16-bit state; the
// last block runs its last repetition at FP32 accuracy from the 16-bit state,
// in the form the in-memory run takes.
//
// ref_mask: illuminate with (u == 1.0f) of the state each chunk visit starts
// from instead of the input mask: the reference's on-disk path re-derives its
// mask that way (store_mask, src/lib/cpp/cpu/diffusion.cc:331-337, :445;
// src/lib/cpp/gpu/diffusion.cc:745-756, :885).  Needs T = 1 and FP32 state.
static int chunked_core(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx, const float *k,
                        int64_t K, int64_t reps, int64_t cz, int64_t T, int precision, void *hs0,
                        void *hs1, uint16_t *out_u16, float *out_f32, int verbose, bool ref_mask) {
    const int64_t R = K / 2, layer = ny * nx, n = nz * layer;
    if (n == 0) return OK;
    const bool lowp = is_lowp(precision), bf = precision == MXB_BF16;
    if (lowp && R > kMaxRadiusLowp) {
        set_error("BF16/FP16 forms support kernel radius <= " + std::to_string(kMaxRadiusLowp));
        return EINVAL_;
    }
    if (ref_mask && (lowp || T != 1)) {
        set_error("the reference on-disk mask rule needs FP32 state and one repetition per chunk visit");
        return EINVAL_;
    }
    const Arith arith = arith_of(lowp ? MXB_FP32 : precision);
    const bool split = precision == MXB_FP32_SPLIT;
    const int64_t esz = lowp ? 2 : 4;  // bytes per host state element
    const int64_t align = lowp || split ? 32 : 1;
    if (T < 1) T = 1;
    if (cz < 1) cz = 1;
    if (cz > nz) cz = nz;
    const int64_t H = T * R;
    const int64_t max_layers = std::min(nz, cz + 2 * H + (align - 1));
    const int64_t cap = max_layers * layer;
    // Two buffer sets on two streams when the device has room for both:
    // consecutive chunks alternate, so one chunk's uploads overlap the
    // previous chunk's compute and downloads.
    struct ChunkSet {
        Stream st;
        DevBuf dm, d0, d1, dq, db;
        hipError_t alloc(int64_t cap, int64_t bits_bytes, bool u16) {
            hipError_t e;
            if ((e = hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking)) != hipSuccess) return e;
            if ((e = dm.alloc(cap)) != hipSuccess || (e = db.alloc(bits_bytes)) != hipSuccess ||
                (e = d0.alloc(state_f32_bytes(cap))) != hipSuccess ||
                (e = d1.alloc(state_f32_bytes(cap))) != hipSuccess)
                return e;
            return u16 ? dq.alloc(cap * 2) : hipSuccess;
        }
    };
    ChunkSet sets[2];
    MXB_HIP(sets[0].alloc(cap, mask_bits_bytes(max_layers, ny, nx), out_u16 != nullptr));
    int nsets = 2;
    if (sets[1].alloc(cap, mask_bits_bytes(max_layers, ny, nx), out_u16 != nullptr) != hipSuccess) {
        (void)hipGetLastError();
        nsets = 1;
    }
    auto sync_all = [&]() -> hipError_t {
        for (int i = 0; i < nsets; ++i) {
            hipError_t e = hipStreamSynchronize(sets[i].st.s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    Stream &st = sets[0].st;
    DevBuf &dm = sets[0].dm, &d0 = sets[0].d0, &dq = sets[0].dq;
    // The host state is first written at the end of the first block: that
    // block builds u0 = (mask > 0) on the device from the mask chunk.
    char *cur = static_cast<char *>(hs0), *nxt = static_cast<char *>(hs1);
    int64_t done = 0;
    if (reps == 0) {
        // u0 chunk by chunk on the device
        for (int64_t z0 = 0; z0 < nz; z0 += cz) {
            const int64_t c = std::min(cz, nz - z0);
            MXB_HIP(hipMemcpyAsync(dm.p, mask + z0 * layer, c * layer, hipMemcpyHostToDevice, st.s));
            MXB_HIP(launch_init_state((const uint8_t *)dm.p, (float *)d0.p, c * layer, st.s));
            MXB_HIP(launch_quantise((float *)d0.p, out_u16 ? (uint16_t *)dq.p : nullptr, nullptr, c * layer, st.s));
            if (out_u16) MXB_HIP(hipMemcpyAsync(out_u16 + z0 * layer, dq.p, c * layer * 2, hipMemcpyDeviceToHost, st.s));
            if (out_f32) MXB_HIP(hipMemcpyAsync(out_f32 + z0 * layer, d0.p, c * layer * 4, hipMemcpyDeviceToHost, st.s));
            MXB_HIP(hipStreamSynchronize(st.s));
        }
        return OK;
    }
    while (done < reps) {
        const int64_t tb = std::min(T, reps - done);
        const bool last_block = done + tb == reps;
        int ci = 0;
        for (int64_t z0 = 0; z0 < nz; z0 += cz, ++ci) {
            ChunkSet &S = sets[ci % nsets];  // its previous chunk's work precedes on its stream
            const hipStream_t cs = S.st.s;
            const int64_t c = std::min(cz, nz - z0);
            const int64_t lo = std::max<int64_t>(0, z0 - tb * R) / align * align;
            const int64_t hi = std::min(nz, z0 + c + tb * R);
            const int64_t lz = hi - lo;
            log_v(verbose, 2, "chunk z0=" + std::to_string(z0) + " layers=" + std::to_string(lz));
            const bool derive_mask = ref_mask && done > 0;
            if (!derive_mask)
                MXB_HIP(hipMemcpyAsync(S.dm.p, mask + lo * layer, lz * layer, hipMemcpyHostToDevice, cs));
            if (done > 0)
                MXB_HIP(hipMemcpyAsync(S.d0.p, cur + lo * layer * esz, lz * layer * esz, hipMemcpyHostToDevice, cs));
            if (derive_mask)
                MXB_HIP(launch_mask_from_state((const float *)S.d0.p, (uint8_t *)S.dm.p, lz * layer, cs));
            MaskBits bits;
            MXB_HIP(pack_bits((const uint8_t *)S.dm.p, lz, ny, nx, (uint32_t *)S.db.p, bits, cs));
            uint16_t *q = last_block && out_u16 ? (uint16_t *)S.dq.p : nullptr;
            float *fin = nullptr;
            uint16_t *fin16 = nullptr;
            int rc;
            if (lowp) {
                rc = run_lowp_core((const uint8_t *)S.dm.p, bits, lz, ny, nx, k, K, last_block ? tb - 1 : tb,
                                   last_block, bf, done == 0, S.d0.p, S.d1.p, q, nullptr, &fin, &fin16, true, cs);
            } else if (split && split_serves(lz, ny, nx, (int)R, S.d0.p, S.d1.p, bits, q, nullptr, true) != 0) {
                // FP32 state between visits, split into planes on the device
                rc = run_steps_split((const uint8_t *)S.dm.p, bits, lz, ny, nx, k, K, tb, S.d0.p, S.d1.p,
                                     done > 0, q, nullptr, &fin,
                                     split_serves(lz, ny, nx, (int)R, S.d0.p, S.d1.p, bits, q, nullptr, true), cs);
            } else {
                if (done == 0)
                    MXB_HIP(launch_init_state((const uint8_t *)S.dm.p, (float *)S.d0.p, lz * layer, cs));
                rc = run_steps_f32((const uint8_t *)S.dm.p, bits, lz, ny, nx, k, K, tb, arith, (float *)S.d0.p,
                                   (float *)S.d1.p, false, q, nullptr, &fin, cs);
            }
            if (rc != OK) {
                (void)sync_all();
                return rc;
            }
            const int64_t off = (z0 - lo) * layer;
            // the interior's state: to the other host buffer, or after the last
            // block to out_f32 (when wanted)
            if (last_block) {
                if (out_f32 != nullptr)
                    MXB_HIP(hipMemcpyAsync(out_f32 + z0 * layer, fin + off, c * layer * 4, hipMemcpyDeviceToHost, cs));
                if (out_u16)
                    MXB_HIP(hipMemcpyAsync(out_u16 + z0 * layer, (uint16_t *)S.dq.p + off, c * layer * 2,
                                           hipMemcpyDeviceToHost, cs));
            } else {
                const char *src = lowp ? reinterpret_cast<const char *>(fin16 + off)
                                       : reinterpret_cast<const char *>(fin + off);
                MXB_HIP(hipMemcpyAsync(nxt + z0 * layer * esz, src, c * layer * esz, hipMemcpyDeviceToHost, cs));
            }
        }
        // the next block reads what this one wrote
        MXB_HIP(sync_all());
        std::swap(cur, nxt);
        done += tb;
        log_v(verbose, 1, "repetition " + std::to_string(done) + "/" + std::to_string(reps));
    }
    return OK;
}

// Host buffer that prefers pinned memory.
struct HostBuf {
    void *p = nullptr;
    bool pinned = false;
    ~HostBuf() {
        if (!p) return;
        if (pinned) (void)hipHostFree(p);
        else std::free(p);
    }
    bool alloc(size_t bytes) {
        if (hipHostMalloc(&p, bytes > 0 ? bytes : 1, hipHostMallocDefault) == hipSuccess) {
            pinned = true;
            return true;
        }
        (void)hipGetLastError();
        p = std::malloc(bytes > 0 ? bytes : 1);
        return p != nullptr;
    }
};

static int chunked_host(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx, const float *k,
                        int64_t K, int64_t reps, int64_t cz, int64_t T, int precision,
                        uint16_t *out_u16, float *out_f32, int verbose) {
    const int64_t n = nz * ny * nx;
    const int64_t esz = is_lowp(precision) ? 2 : 4;
    HostBuf h0, h1;
    // host state carries the volume between blocks of T repetitions; one
    // block (reps <= T) needs none
    if (reps > std::max<int64_t>(T, 1) && (!h0.alloc(n * esz) || !h1.alloc(n * esz))) {
        set_error("host allocation of the chunked state failed");
        return ENOMEM_;
    }
    return chunked_core(mask, nz, ny, nx, k, K, reps, cz, T, precision, h0.p, h1.p, out_u16, out_f32,
                        verbose, false);
}

// ---------------------------------------------------------------- staged copies
// Copies between pageable host memory and the device through two pinned
// 64 MB chunks: host threads fill (or drain) one chunk while the DMA engine
// moves the other.  A pageable hipMemcpy runs at ~10 GB/s on the MI355X
// boxes; the staged form moves several times that.  The pinned chunks are
// allocated at the first large copy and kept until the process ends; one
// staged copy runs at a time (a mutex).  Small copies, or a failed pinned
// allocation, take the plain path.
namespace {
constexpr size_t kStageChunk = size_t(64) << 20;
std::mutex g_stage_mu;
void *g_stage[2] = {nullptr, nullptr};
hipEvent_t g_stage_ev[2] = {nullptr, nullptr};

void parallel_memcpy(void *dst, const void *src, size_t bytes) {
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt == 0 ? 4 : std::min(nt, 16u);
    const size_t part = (bytes + nt - 1) / nt;
    std::vector<std::thread> th;
    for (unsigned i = 0; i < nt; ++i) {
        const size_t b = i * part;
        if (b >= bytes) break;
        const size_t len = std::min(part, bytes - b);
        th.emplace_back([=] { std::memcpy(static_cast<char *>(dst) + b, static_cast<const char *>(src) + b, len); });
    }
    for (auto &t : th) t.join();
}

bool stage_ready() {
    if (g_stage[0] != nullptr) return true;
    for (int i = 0; i < 2; ++i) {
        if (hipHostMalloc(&g_stage[i], kStageChunk, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&g_stage_ev[i], hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            for (int j = 0; j <= i; ++j) {
                if (g_stage[j]) (void)hipHostFree(g_stage[j]);
                if (g_stage_ev[j]) (void)hipEventDestroy(g_stage_ev[j]);
                g_stage[j] = nullptr;
                g_stage_ev[j] = nullptr;
            }
            return false;
        }
    }
    return true;
}
}  // namespace

// Host -> device (h2d) or device -> host copy of `bytes`, ordered on `s`
// after the work already queued there; returns with the copy complete.
hipError_t staged_copy(void *dst, const void *src, size_t bytes, bool h2d, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    if (bytes < 2 * kStageChunk || !stage_ready()) {
        hipError_t e = hipMemcpyAsync(dst, src, bytes, h2d ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost, s);
        return e != hipSuccess ? e : hipStreamSynchronize(s);
    }
    const size_t nch = (bytes + kStageChunk - 1) / kStageChunk;
    auto len_of = [&](size_t i) { return std::min(kStageChunk, bytes - i * kStageChunk); };
    hipError_t e;
    if (h2d) {
        for (size_t i = 0; i < nch; ++i) {
            const int b = (int)(i % 2);
            if (i >= 2 && (e = hipEventSynchronize(g_stage_ev[b])) != hipSuccess) return e;
            parallel_memcpy(g_stage[b], static_cast<const char *>(src) + i * kStageChunk, len_of(i));
            if ((e = hipMemcpyAsync(static_cast<char *>(dst) + i * kStageChunk, g_stage[b], len_of(i),
                                    hipMemcpyHostToDevice, s)) != hipSuccess ||
                (e = hipEventRecord(g_stage_ev[b], s)) != hipSuccess)
                return e;
        }
        return hipStreamSynchronize(s);
    }
    auto enqueue = [&](size_t i) -> hipError_t {
        const int b = (int)(i % 2);
        hipError_t r = hipMemcpyAsync(g_stage[b], static_cast<const char *>(src) + i * kStageChunk, len_of(i),
                                      hipMemcpyDeviceToHost, s);
        return r != hipSuccess ? r : hipEventRecord(g_stage_ev[b], s);
    };
    for (size_t i = 0; i < std::min<size_t>(2, nch); ++i)
        if ((e = enqueue(i)) != hipSuccess) return e;
    for (size_t i = 0; i < nch; ++i) {
        const int b = (int)(i % 2);
        if ((e = hipEventSynchronize(g_stage_ev[b])) != hipSuccess) return e;
        parallel_memcpy(static_cast<char *>(dst) + i * kStageChunk, g_stage[b], len_of(i));
        if (i + 2 < nch && (e = enqueue(i + 2)) != hipSuccess) return e;
    }
    return hipSuccess;
}

// Device memory the drivers may plan with: 92 % of what is free, less
// 256 MB.  MXB_DEVICE_BUDGET (bytes) lowers it, so that small volumes can be
// sent down the streaming paths that volumes larger than HBM take.
static int device_budget(int64_t &budget) {
    size_t free_b = 0, total_b = 0;
    MXB_HIP(hipMemGetInfo(&free_b, &total_b));
    budget = (int64_t)(free_b * 0.92) - (int64_t(256) << 20);
    if (const char *v = std::getenv("MXB_DEVICE_BUDGET")) {
        char *end = nullptr;
        const long long cap = std::strtoll(v, &end, 10);
        if (end != v && cap >= 0) budget = std::min<int64_t>(budget, cap);
    }
    return OK;
}

// Largest chunk depth whose device footprint fits in `budget` bytes.
static int64_t chunk_depth_for(int64_t budget, int64_t nz, int64_t layer, int64_t H) {
    const int64_t per_layer = layer * (1 + 4 + 4 + 2);
    int64_t layers = budget / std::max<int64_t>(per_layer, 1);
    int64_t cz = layers - 2 * H;
    if (cz < 16) cz = 16;
    return std::min(cz, nz);
}

static int in_memory(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx, const float *k,
                     int64_t K, int64_t reps, int precision, uint16_t *out_u16, float *out_f32,
                     int verbose) {
    int rc = check_common(nz, ny, nx, k, K, reps, precision);
    if (rc != OK) return rc;
    if (out_u16 == nullptr && out_f32 == nullptr) {
        set_error("no output buffer given");
        return EINVAL_;
    }
    const int64_t n = nz * ny * nx;
    if (n == 0) return OK;
    if (mask == nullptr) {
        set_error("mask is null");
        return EINVAL_;
    }
    int64_t budget = 0;
    rc = device_budget(budget);
    if (rc != OK) return rc;
    const int64_t sb = state_bytes(precision, nz, ny, nx);
    const int64_t need = n + 2 * sb + (out_u16 ? 2 * n : 0) + (out_f32 ? 4 * n : 0);
    if (need > budget) {
        const int64_t H = K / 2;
        const int64_t cz = chunk_depth_for(budget, nz, ny * nx, H);
        log_v(verbose, 1, "volume exceeds device memory: streaming z-chunks of " + std::to_string(cz));
        return chunked_host(mask, nz, ny, nx, k, K, reps, cz, 1, precision, out_u16, out_f32, verbose);
    }
    Stream st;
    MXB_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    DevBuf dm, d0, d1, dq, df;
    MXB_HIP(dm.alloc(n));
    MXB_HIP(d0.alloc(sb));
    MXB_HIP(d1.alloc(sb));
    if (out_u16) MXB_HIP(dq.alloc(n * 2));
    if (out_f32) MXB_HIP(df.alloc(n * 4));
    log_v(verbose, 1, "Converting uint8 to float");
    MXB_HIP(staged_copy(dm.p, mask, n, true, st.s));
    rc = run_steps((const uint8_t *)dm.p, nz, ny, nx, k, K, reps, precision, d0.p, d1.p,
                   (uint16_t *)dq.p, (float *)df.p, st.s);
    if (rc != OK) return rc;
    if (out_u16) MXB_HIP(staged_copy(out_u16, dq.p, n * 2, false, st.s));
    if (out_f32) MXB_HIP(staged_copy(out_f32, df.p, n * 4, false, st.s));
    MXB_HIP(hipStreamSynchronize(st.s));
    log_v(verbose, 1, "Diffusion is complete!");
    return OK;
}

// ---------------------------------------------------------------- files
struct MappedFile {
    int fd = -1;
    void *p = MAP_FAILED;
    size_t len = 0;
    ~MappedFile() {
        if (p != MAP_FAILED && len > 0) munmap(p, len);
        if (fd >= 0) close(fd);
    }
};

static int map_read(const char *path, size_t need, MappedFile &m) {
    m.fd = open(path, O_RDONLY);
    if (m.fd < 0) {
        set_error(std::string("Could not open file: ") + path);
        return EIO_;
    }
    struct stat sb;
    if (fstat(m.fd, &sb) != 0 || (size_t)sb.st_size < need) {
        set_error(std::string("Could not read the entire buffer: ") + path + " holds fewer than " +
                  std::to_string(need) + " bytes");
        return EIO_;
    }
    m.len = need;
    if (need == 0) return OK;
    m.p = mmap(nullptr, need, PROT_READ, MAP_SHARED, m.fd, 0);
    if (m.p == MAP_FAILED) {
        set_error(std::string("mmap failed: ") + path);
        return EIO_;
    }
    return OK;
}

static int map_write(const char *path, size_t len, MappedFile &m, bool unlink_after) {
    if (path != nullptr) {
        m.fd = open(path, O_RDWR | O_CREAT | O_TRUNC, 0644);
    } else {
        const char *tmpdir = std::getenv("TMPDIR");
        std::string tmpl = std::string(tmpdir && *tmpdir ? tmpdir : "/tmp") + "/mxb-diffusion-XXXXXX";
        std::vector<char> buf(tmpl.begin(), tmpl.end());
        buf.push_back('\0');
        m.fd = mkstemp(buf.data());
        if (m.fd >= 0 && unlink_after) unlink(buf.data());
    }
    if (m.fd < 0) {
        set_error(std::string("Could not open file: ") + (path ? path : "temporary state file"));
        return EIO_;
    }
    if (ftruncate(m.fd, (off_t)len) != 0) {
        set_error(std::string("Could not truncate file: ") + (path ? path : "temporary state file"));
        return EIO_;
    }
    m.len = len;
    if (len == 0) return OK;
    m.p = mmap(nullptr, len, PROT_READ | PROT_WRITE, MAP_SHARED, m.fd, 0);
    if (m.p == MAP_FAILED) {
        set_error(std::string("mmap failed: ") + (path ? path : "temporary state file"));
        return EIO_;
    }
    return OK;
}

// ---------------------------------------------------------------- no-allocation entry
// The two state buffers of mxb_device_diffusion_noalloc: a file-scope device
// array that every device the library runs on holds until the process ends.
// In host code `scratch` only names the symbol; its address on a device comes
// from hipGetSymbolAddress with that device current, looked up once per
// device and kept in g_scratch_ptr.
static __device__ __align__(256) unsigned char scratch[MXB_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

struct TimingOff {
    TimingOff() { g_timing_off = true; }
    ~TimingOff() { g_timing_off = false; }
};
}  // namespace

// Device address of `scratch` on the current device.
static int scratch_address(void **out) {
    int dev = 0;
    MXB_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device ordinal " + std::to_string(dev) + " beyond the supported " + std::to_string(kMaxDevices));
        return EINVAL_;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        MXB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) {
            set_error("hipGetSymbolAddress returned no address for the scratch array");
            return EDEVICE;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return OK;
}

}  // namespace mxb

using namespace mxb;

extern "C" {

const char *mxb_last_error(void) { return g_error.c_str(); }
const char *mxb_version(void) { return "maxibone_amd diffusion 0.1 (gfx950)"; }

int64_t mxb_state_bytes(int precision, int64_t nz, int64_t ny, int64_t nx) {
    if (precision < MXB_FP32 || precision > MXB_FP32_SPLIT || nz < 0 || ny < 0 || nx < 0) {
        set_error("mxb_state_bytes: unknown precision or negative extent");
        return -1;
    }
    return state_bytes(precision, nz, ny, nx);
}

int mxb_diffusion_ex(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx, const float *kernel,
                     int64_t kernel_size, int64_t repetitions, int precision, uint16_t *output_u16,
                     float *output_f32, int verbose) {
    return in_memory(mask, nz, ny, nx, kernel, kernel_size, repetitions, precision, output_u16,
                     output_f32, verbose);
}

int mxb_diffusion_u8_u16(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx,
                         const float *kernel, int64_t kernel_size, int64_t repetitions,
                         uint16_t *output, int verbose) {
    if (output == nullptr) {
        set_error("output is null");
        return EINVAL_;
    }
    return in_memory(mask, nz, ny, nx, kernel, kernel_size, repetitions, MXB_FP32, output, nullptr,
                     verbose);
}

int mxb_diffusion_u8_f32(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx,
                         const float *kernel, int64_t kernel_size, int64_t repetitions,
                         float *output, int verbose) {
    if (output == nullptr) {
        set_error("output is null");
        return EINVAL_;
    }
    return in_memory(mask, nz, ny, nx, kernel, kernel_size, repetitions, MXB_FP32, nullptr, output,
                     verbose);
}

int mxb_diffusion_chunked(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx,
                          const float *kernel, int64_t kernel_size, int64_t repetitions,
                          int64_t chunk_z, int64_t temporal, int precision, uint16_t *output_u16,
                          float *output_f32, int verbose) {
    int rc = check_common(nz, ny, nx, kernel, kernel_size, repetitions, precision);
    if (rc != OK) return rc;
    if (output_u16 == nullptr && output_f32 == nullptr) {
        set_error("no output buffer given");
        return EINVAL_;
    }
    if (chunk_z < 1) {
        set_error("chunk_z must be positive");
        return EINVAL_;
    }
    if (temporal < 1) {
        set_error("temporal must be positive");
        return EINVAL_;
    }
    if (nz * ny * nx > 0 && mask == nullptr) {
        set_error("mask is null");
        return EINVAL_;
    }
    return chunked_host(mask, nz, ny, nx, kernel, kernel_size, repetitions, chunk_z, temporal,
                        precision, output_u16, output_f32, verbose);
}

int mxb_diffusion_on_disk_ex(const char *input_file, const float *kernel, int64_t kernel_size,
                             const char *output_file, int64_t total_z, int64_t total_y,
                             int64_t total_x, int64_t global_z, int64_t global_y, int64_t global_x,
                             int64_t repetitions, int precision, int flags, int verbose) {
    int rc = check_common(total_z, total_y, total_x, kernel, kernel_size, repetitions, precision);
    if (rc != OK) return rc;
    if (input_file == nullptr || output_file == nullptr) {
        set_error("file names must not be null");
        return EINVAL_;
    }
    if (global_z < 1 || global_y != total_y || global_x != total_x) {
        set_error("global_shape must be (chunk_z >= 1, total_y, total_x): chunks split z only");
        return EINVAL_;
    }
    if ((flags & ~MXB_ON_DISK_REFERENCE_MASK) != 0) {
        set_error("unknown on-disk flags " + std::to_string(flags));
        return EINVAL_;
    }
    const bool ref_mask = (flags & MXB_ON_DISK_REFERENCE_MASK) != 0;
    if (ref_mask && is_lowp(precision)) {
        set_error("the reference on-disk mask rule (u == 1.0f) needs an FP32 precision");
        return EINVAL_;
    }
    const int64_t n = total_z * total_y * total_x;
    const int64_t esz = is_lowp(precision) ? 2 : 4;
    MappedFile in, out;
    rc = map_read(input_file, (size_t)n, in);
    if (rc != OK) return rc;
    rc = map_write(output_file, (size_t)n * 2, out, false);
    if (rc != OK) return rc;
    if (n == 0) return OK;
    // State lives in two unlinked temporary files, paged by the OS (unique
    // names, so concurrent calls do not collide).
    MappedFile t0, t1;
    rc = map_write(nullptr, (size_t)(n * esz), t0, true);
    if (rc != OK) return rc;
    rc = map_write(nullptr, (size_t)(n * esz), t1, true);
    if (rc != OK) return rc;
    int64_t budget = 0;
    rc = device_budget(budget);
    if (rc != OK) return rc;
    const int64_t R = kernel_size / 2, layer = total_y * total_x;
    // The caller's chunk depth is honoured.  Temporal blocking runs T
    // repetitions per chunk visit (halo T*R layers per side), which divides
    // the file traffic by T; T is capped so that the halo stays within about
    // one chunk depth (at most ~3x redundant work) and the chunk fits.  The
    // reference mask rule re-derives the mask per visit, so it keeps T = 1.
    const int64_t pad = is_lowp(precision) ? 31 : 0;  // the 16-bit forms align chunk starts to 32 layers
    int64_t T = 1;
    while (!ref_mask && T < repetitions && 2 * (T + 1) * R <= std::max(global_z, R) &&
           std::min(total_z, global_z + 2 * (T + 1) * R + pad) * layer * 11 < budget)
        ++T;
    if (std::min(total_z, global_z + 2 * T * R + pad) * layer * 11 > budget) {
        set_error("chunk of " + std::to_string(global_z) + " layers plus halo exceeds device memory");
        return ENOMEM_;
    }
    rc = chunked_core((const uint8_t *)in.p, total_z, total_y, total_x, kernel, kernel_size,
                      repetitions, global_z, T, precision, t0.p, t1.p, (uint16_t *)out.p, nullptr,
                      verbose, ref_mask);
    if (rc != OK) return rc;
    if (msync(out.p, out.len, MS_SYNC) != 0) {
        set_error(std::string("Could not write the output file: ") + output_file);
        return EIO_;
    }
    return OK;
}

int mxb_diffusion_on_disk(const char *input_file, const float *kernel, int64_t kernel_size,
                          const char *output_file, int64_t total_z, int64_t total_y,
                          int64_t total_x, int64_t global_z, int64_t global_y, int64_t global_x,
                          int64_t repetitions, int verbose) {
    return mxb_diffusion_on_disk_ex(input_file, kernel, kernel_size, output_file, total_z, total_y,
                                    total_x, global_z, global_y, global_x, repetitions, MXB_FP32, 0,
                                    verbose);
}

int mxb_device_diffusion(const uint8_t *dev_mask, int64_t nz, int64_t ny, int64_t nx,
                         const float *kernel, int64_t kernel_size, int64_t repetitions,
                         int precision, void *dev_state0, void *dev_state1, uint16_t *dev_out_u16,
                         float *dev_out_f32, void *stream) {
    return device_diffusion(dev_mask, nz, ny, nx, kernel, kernel_size, repetitions, precision, dev_state0,
                            dev_state1, dev_out_u16, dev_out_f32, stream, true);
}

}  // extern "C"

namespace mxb {

int device_diffusion(const uint8_t *dev_mask, int64_t nz, int64_t ny, int64_t nx, const float *kernel,
                     int64_t kernel_size, int64_t repetitions, int precision, void *dev_state0,
                     void *dev_state1, uint16_t *dev_out_u16, float *dev_out_f32, void *stream, bool fused) {
    int rc = check_common(nz, ny, nx, kernel, kernel_size, repetitions, precision);
    if (rc != OK) return rc;
    if (dev_out_u16 == nullptr && dev_out_f32 == nullptr) {
        set_error("no output buffer given");
        return EINVAL_;
    }
    if (nz * ny * nx > 0 && (dev_mask == nullptr || dev_state0 == nullptr || dev_state1 == nullptr)) {
        set_error("device buffers must not be null");
        return EINVAL_;
    }
    return run_steps(dev_mask, nz, ny, nx, kernel, kernel_size, repetitions, precision, dev_state0,
                     dev_state1, dev_out_u16, dev_out_f32, (hipStream_t)stream, fused);
}

}  // namespace mxb

extern "C" {

int mxb_device_diffusion_noalloc(const uint8_t *dev_mask, int64_t nz, int64_t ny, int64_t nx,
                                 const float *kernel, int64_t kernel_size, int64_t repetitions,
                                 int precision, uint16_t *dev_out_u16, float *dev_out_f32,
                                 void *stream) {
    int rc = check_common(nz, ny, nx, kernel, kernel_size, repetitions, precision);
    if (rc != OK) return rc;
    if (dev_out_u16 == nullptr && dev_out_f32 == nullptr) {
        set_error("no output buffer given");
        return EINVAL_;
    }
    if (nz * ny * nx == 0) return OK;
    if (dev_mask == nullptr) {
        set_error("device mask must not be null");
        return EINVAL_;
    }
    // two state buffers, the second at the next 256-byte boundary
    const int64_t sb = state_bytes(precision, nz, ny, nx);
    const int64_t off = (sb + 255) / 256 * 256;
    if (off + sb > (int64_t)MXB_NOALLOC_SCRATCH_BYTES) {
        set_error("volume needs " + std::to_string(off + sb) + " bytes of state; the no-allocation entry "
                  "holds at most " + std::to_string((int64_t)MXB_NOALLOC_SCRATCH_BYTES));
        return EINVAL_;
    }
    void *base = nullptr;
    rc = scratch_address(&base);
    if (rc != OK) return rc;
    // Every byte the kernels read from the state buffers is written earlier in
    // the same call (initial state, packed mask bits, pass outputs), so
    // nothing needs clearing between calls.
    TimingOff no_events;
    return run_steps(dev_mask, nz, ny, nx, kernel, kernel_size, repetitions, precision, base,
                     static_cast<char *>(base) + off, dev_out_u16, dev_out_f32, (hipStream_t)stream);
}

int mxb_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int mxb_set_kernel_timing(int enable) {
    std::lock_guard<std::mutex> lk(g_timing.mu);
    drain_timing_locked();
    for (int i = 0; i < 5; ++i) {
        g_timing.total_ms[i] = 0;
        g_timing.launches[i] = 0;
    }
    g_timing.enabled = enable != 0;
    return OK;
}

int mxb_kernel_times(int kernel_class, double *total_ms, int64_t *launches) {
    if (kernel_class < 0 || kernel_class > MXB_K_XYPASS) {
        set_error("unknown kernel class");
        return EINVAL_;
    }
    std::lock_guard<std::mutex> lk(g_timing.mu);
    drain_timing_locked();
    if (total_ms) *total_ms = g_timing.total_ms[kernel_class];
    if (launches) *launches = g_timing.launches[kernel_class];
    return OK;
}

}  // extern "C"
