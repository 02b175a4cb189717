// Batched diffusion: many independent sub-volumes of uneven shape and radius.
//
// One persistent kernel per pass serves every volume: the work list holds one
// item per (volume, 64-line group, outer index, segment), sorted by cost
// (tiles x MFMA chunks) so that a grid-stride walk over it balances the
// blocks; each item runs the direct-load line-pass body (mxb_f32_dma_body.h)
// compiled for the smallest radius bucket that holds the volume's radius.
// Each volume's state is kept in a pitched layout, (z, py, px) for the
// (z, y, x) order and (z, px, py) for the (z, x, y) order, with px, py the
// extents rounded up to 128 bytes of state, so every pass meets the body's
// alignment and whole-piece requirements for any shape.  The passes run over the padded
// line counts (px or py lines); pad lines only ever feed pad outputs (each
// line is independent and no pass reads positions beyond the true extent),
// so the valid region does not depend on what the pads hold.  The mask is
// copied into the same pitch (zero pads), and a final per-volume kernel
// writes the caller's unpadded outputs.
// The whole repetition loop is captured once in a HIP graph; replays launch it.
//
// The grid size and the segment length are chosen by timing candidates on
// repetitions of the actual batch (autotune, the fastest of three timings
// each); the choice is cached per batch signature in the process (and in the
// file named by MXB_TUNE_FILE, if set).
//
// BF16 / FP16 plans run the 16-bit direct-load body (mxb_lowp_dma_body.h) in
// the same way for repetitions 1 .. reps-1, with the 16-bit state in the first
// half of each volume's pitched state, then widen it and run the last
// repetition with the FP32 kernels, as the single-volume path does.  Their
// radius buckets are the single-volume 16-bit ones (the bucket fixes how taps
// group into 32-position MFMA chunks), so the results are bit-identical to it.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "maxibone_diffusion.h"
#include "mxb_f32_dma_body.h"
#include "mxb_lowp_dma_body.h"

namespace mxb {

// Radius buckets of the persistent kernel (compiled tap capacities).
static constexpr int kBucketRT[] = {8, 12, 16, 20, 24, 32, 40, 48, 64};
static constexpr int kNumBuckets = sizeof(kBucketRT) / sizeof(kBucketRT[0]);
// Prefetch depths of the 16-bit persistent kernel (tile pairs ahead): the
// z form, and the y and x forms (the masked body needs D - 1 to be a
// multiple of 4).
#ifndef MXB_LOWP_BATCH_DEPTH_Z
#define MXB_LOWP_BATCH_DEPTH_Z 5
#endif
#ifndef MXB_LOWP_BATCH_DEPTH
#define MXB_LOWP_BATCH_DEPTH 5
#endif
// Row pitch of the padded layout: 128 bytes of state elements (one L2 line),
// so that the 64-line rows the passes move start on a line boundary and the
// line groups of neighbouring blocks never share a line (a 16-element pitch
// of 16-bit state left the rows 32-byte aligned: 0.69x the throughput).
static int64_t pitch_elements(int precision) { return precision == MXB_BF16 || precision == MXB_FP16 ? 64 : 32; }
template <int FORM>
static constexpr int kLowpBatchDepth = FORM == 0 ? MXB_LOWP_BATCH_DEPTH_Z : MXB_LOWP_BATCH_DEPTH;
// Radius buckets of the 16-bit persistent kernel: those of the single-volume
// 16-bit passes (launch_line_pass_lowp_lds).
static constexpr int kLowpBucketRT[] = {4, 8, 12, 16, 24, 32, 40, 48, 64};
static constexpr int kNumLowpBuckets = sizeof(kLowpBucketRT) / sizeof(kLowpBucketRT[0]);

struct BatchPassDesc {
    const float *in;
    LineGeom g;
    Epilogue e;
    int32_t bucket;
    int32_t taps_off;  // offset of 2*RT+1 centred taps in the tap table
};

struct BatchItem {
    int32_t vol, seg, lb, outer;
};

// Prefetch depths of the FP32 persistent kernel: the z form (tiles ahead) and
// the y and x forms (tile pairs ahead; the masked body needs D - 1 to be a
// multiple of 4).
#ifndef MXB_F32_BATCH_DEPTH_Z
#define MXB_F32_BATCH_DEPTH_Z 2
#endif
#ifndef MXB_F32_BATCH_DEPTH_YX
#define MXB_F32_BATCH_DEPTH_YX 1
#endif

// Pass forms of the persistent kernel: 0 z (line-contiguous output), 1 y,
// 2 x (position-contiguous output; x illuminates).
template <int FORM, int RT>
__device__ __forceinline__ void batch_body(const BatchPassDesc &d, const BatchItem &it,
                                           const float *__restrict__ taps, unsigned char *smem) {
    if constexpr (FORM == 0)
        f32_dma_body<RT, MXB_F32_BATCH_DEPTH_Z, false, true, false, true>(d.in, d.g, d.e, it.outer, it.lb, it.seg,
                                                                          taps + d.taps_off, smem);
    else
        f32_dma_body2<RT, MXB_F32_BATCH_DEPTH_YX, FORM == 2, true, false, false>(d.in, d.g, d.e, it.outer, it.lb,
                                                                                 it.seg, taps + d.taps_off, smem);
}

template <int FORM>
__global__ __launch_bounds__(256) void k_batched_pass(const BatchPassDesc *__restrict__ descs,
                                                      const BatchItem *__restrict__ items,
                                                      int n_items,
                                                      const float *__restrict__ taps) {
    extern __shared__ __attribute__((aligned(16))) unsigned char bsmem[];
    for (int i = blockIdx.x; i < n_items; i += gridDim.x) {
        const BatchItem it = items[i];
        const BatchPassDesc d = descs[it.vol];
        switch (d.bucket) {
            case 0: batch_body<FORM, 8>(d, it, taps, bsmem); break;
            case 1: batch_body<FORM, 12>(d, it, taps, bsmem); break;
            case 2: batch_body<FORM, 16>(d, it, taps, bsmem); break;
            case 3: batch_body<FORM, 20>(d, it, taps, bsmem); break;
            case 4: batch_body<FORM, 24>(d, it, taps, bsmem); break;
            case 5: batch_body<FORM, 32>(d, it, taps, bsmem); break;
            case 6: batch_body<FORM, 40>(d, it, taps, bsmem); break;
            case 7: batch_body<FORM, 48>(d, it, taps, bsmem); break;
            default: batch_body<FORM, 64>(d, it, taps, bsmem); break;
        }
    }
}

template <int FORM, int RT>
static constexpr int form_smem() {
    if constexpr (FORM == 0)
        return F32DmaSmem<RT, MXB_F32_BATCH_DEPTH_Z, false, true>::BYTES;
    else
        return F32Dma2Smem<RT, MXB_F32_BATCH_DEPTH_YX, FORM == 2, false>::BYTES;
}

template <int FORM>
static int smem_bytes_of(int bucket) {
    switch (bucket) {
        case 0: return form_smem<FORM, 8>();
        case 1: return form_smem<FORM, 12>();
        case 2: return form_smem<FORM, 16>();
        case 3: return form_smem<FORM, 20>();
        case 4: return form_smem<FORM, 24>();
        case 5: return form_smem<FORM, 32>();
        case 6: return form_smem<FORM, 40>();
        case 7: return form_smem<FORM, 48>();
        default: return form_smem<FORM, 64>();
    }
}

static int chunks_of(int bucket) { return (16 + 2 * kBucketRT[bucket] + 3) / 4; }

// ---- 16-bit passes
struct BatchLowpDesc {
    const uint16_t *in;
    uint16_t *out;
    const uint8_t *mask;  // x-pass: pitched byte mask (partial tiles)
    MaskBits bits;        // x-pass: pitched mask bits
    LineGeom g;
    int32_t bucket;       // index into kLowpBucketRT
    int32_t taps_off;     // 2 RT + 1 high parts, then 2 RT + 1 low parts (uint16)
};

template <int FORM, bool BF, int RT>
__device__ __forceinline__ void batch_lowp_body(const BatchLowpDesc &d, const BatchItem &it,
                                                const uint16_t *__restrict__ taps, unsigned char *smem) {
    lowp_dma_body<RT, kLowpBatchDepth<FORM>, BF, FORM == 2, FORM == 0>(
        d.in, d.out, d.mask, d.bits, d.g, taps + d.taps_off, taps + d.taps_off + 2 * RT + 1, it.outer, it.lb,
        it.seg, smem);
}

template <int FORM, bool BF>
__global__ __launch_bounds__(256) void k_batched_lowp_pass(const BatchLowpDesc *__restrict__ descs,
                                                           const BatchItem *__restrict__ items, int n_items,
                                                           const uint16_t *__restrict__ taps) {
    extern __shared__ __attribute__((aligned(16))) unsigned char bsmem[];
    for (int i = blockIdx.x; i < n_items; i += gridDim.x) {
        const BatchItem it = items[i];
        const BatchLowpDesc d = descs[it.vol];
        switch (d.bucket) {
            case 0: batch_lowp_body<FORM, BF, 4>(d, it, taps, bsmem); break;
            case 1: batch_lowp_body<FORM, BF, 8>(d, it, taps, bsmem); break;
            case 2: batch_lowp_body<FORM, BF, 12>(d, it, taps, bsmem); break;
            case 3: batch_lowp_body<FORM, BF, 16>(d, it, taps, bsmem); break;
            case 4: batch_lowp_body<FORM, BF, 24>(d, it, taps, bsmem); break;
            case 5: batch_lowp_body<FORM, BF, 32>(d, it, taps, bsmem); break;
            case 6: batch_lowp_body<FORM, BF, 40>(d, it, taps, bsmem); break;
            case 7: batch_lowp_body<FORM, BF, 48>(d, it, taps, bsmem); break;
            default: batch_lowp_body<FORM, BF, 64>(d, it, taps, bsmem); break;
        }
    }
}

template <int FORM, int RT>
static constexpr int lowp_form_smem() {
    return LowpDmaSmem<RT, kLowpBatchDepth<FORM>, FORM == 2, FORM == 0>::BYTES;
}

template <int FORM>
static int lowp_smem_bytes_of(int bucket) {
    switch (bucket) {
        case 0: return lowp_form_smem<FORM, 4>();
        case 1: return lowp_form_smem<FORM, 8>();
        case 2: return lowp_form_smem<FORM, 12>();
        case 3: return lowp_form_smem<FORM, 16>();
        case 4: return lowp_form_smem<FORM, 24>();
        case 5: return lowp_form_smem<FORM, 32>();
        case 6: return lowp_form_smem<FORM, 40>();
        case 7: return lowp_form_smem<FORM, 48>();
        default: return lowp_form_smem<FORM, 64>();
    }
}

// The centred 16-bit tap split of split_taps<RT> (mxb_lowp.h) for a runtime RT.
static void split_taps_into(const float *k, int R, int RT, bool bf, uint16_t *hi, uint16_t *lo) {
    for (int i = 0; i < 2 * RT + 1; ++i) {
        const int r = i - RT;
        const float v = (r >= -R && r <= R) ? k[R + r] : 0.0f;
        const uint16_t h = bf ? bf16_bits(v) : f16_bits(v);
        const float rest = v - (bf ? bf16_value(h) : f16_value(h));
        hi[i] = h;
        lo[i] = bf ? bf16_bits(rest) : f16_bits(rest);
    }
}

// Pitched-layout initialisation and final unpadding.
__global__ __launch_bounds__(256) void k_unpad_out(const float *__restrict__ st, int64_t nz,
                                                   int64_t ny, int64_t nx, int64_t py, int64_t px,
                                                   uint16_t *__restrict__ q, float *__restrict__ f) {
    const int64_t n = nz * ny * nx;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t x = i % nx, zy = i / nx;
        const int64_t y = zy % ny, z = zy / ny;
        const float v = st[(z * py + y) * px + x];
        if (q != nullptr) q[i] = quantise_u16_l(v);
        if (f != nullptr) f[i] = v;
    }
}

// Copy a (nz, ny, nx) mask into the (nz, py, px) pitch (pads stay as set).
__global__ __launch_bounds__(256) void k_pitch_mask(const uint8_t *__restrict__ m, int64_t nz,
                                                    int64_t ny, int64_t nx, int64_t py, int64_t px,
                                                    uint8_t *__restrict__ dst) {
    const int64_t n = nz * ny * nx;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t x = i % nx, zy = i / nx;
        const int64_t y = zy % ny, z = zy / ny;
        dst[(z * py + y) * px + x] = m[i];
    }
}

__global__ __launch_bounds__(256) void k_init_pitched(const uint8_t *__restrict__ m, float *__restrict__ u,
                                                      int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        u[i] = m[i] > 0 ? 1.0f : 0.0f;
}

static unsigned blocks_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

// ----------------------------------------------------------------- the plan
struct VolLayout {
    int64_t nz, ny, nx, px, py;
    int64_t st_off;    // floats, into each state buffer
    int64_t mask_off;  // bytes, into the pitched masks
    int64_t bits_off;  // words, into the pitched mask bits (nz x words_per_row(px) x py)
    int R, bucket;
    int lbucket;       // 16-bit radius bucket
};

struct TuneChoice {
    int grid = 0;
    int seg = 0;  // 0: whole lines
    double ms = 0;
};

static std::mutex g_tune_mu;
static std::map<std::string, TuneChoice> g_tune;

static void load_tune_file() {
    static bool loaded = false;
    if (loaded) return;
    loaded = true;
    const char *path = std::getenv("MXB_TUNE_FILE");
    if (!path) return;
    FILE *f = std::fopen(path, "r");
    if (!f) return;
    char key[512];
    TuneChoice c;
    while (std::fscanf(f, "%511s %d %d %lf", key, &c.grid, &c.seg, &c.ms) == 4) g_tune[key] = c;
    std::fclose(f);
}

static void save_tune_file() {
    const char *path = std::getenv("MXB_TUNE_FILE");
    if (!path) return;
    FILE *f = std::fopen(path, "w");
    if (!f) return;
    for (auto &kv : g_tune) std::fprintf(f, "%s %d %d %.6f\n", kv.first.c_str(), kv.second.grid, kv.second.seg, kv.second.ms);
    std::fclose(f);
}

}  // namespace mxb

struct mxb_batch_plan {
    int n = 0;
    int64_t reps = 0;
    int precision = MXB_FP32;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<mxb_device_volume> vols;
    std::vector<mxb::VolLayout> lay;
    std::vector<std::vector<float>> host_taps;
    void *ws = nullptr;  // one device workspace
    uint8_t *masks = nullptr;
    uint32_t *bits = nullptr;
    float *st[2] = {nullptr, nullptr};
    float *taps = nullptr;
    mxb::BatchPassDesc *descs = nullptr;  // [6][n]: (parity, pass)
    mxb::BatchLowpDesc *ldescs = nullptr; // [6][n]: (parity, pass), 16-bit plans
    uint16_t *ltaps = nullptr;            // 16-bit tap table
    bool lowp = false;                    // BF16 / FP16 plan with the persistent form
    int lsmem_bytes[3] = {0, 0, 0};
    mxb::BatchItem *items = nullptr;      // z | y | x lists
    int n_items[3] = {0, 0, 0};
    int64_t item_off[3] = {0, 0, 0};
    int item_cap = 0;
    int smem_bytes[3] = {0, 0, 0};  // per pass form
    mxb::TuneChoice tune;
    bool persistent = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string key;
};

namespace mxb {

#define MXB_BHIP(expr)                                                                  \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess) {                                                         \
            set_error(std::string(#expr) + ": " + hipGetErrorString(_e));               \
            return _e == hipErrorOutOfMemory ? ENOMEM_ : EDEVICE;                       \
        }                                                                               \
    } while (0)

// Geometry of pass `p` (0 z, 1 y, 2 x) of a pitched volume: (z, py, px)
// order in and out of z, (z, px, py) between y and x.  Lines run over the
// padded counts; positions over the true extents; the z-pass skips pad rows.
static LineGeom pitched_geom(const VolLayout &v, int p) {
    LineGeom g{};
    if (p == 0) {
        g.n_lane = v.px; g.n_outer = v.ny; g.n_ax = v.nz;
        g.in_outer = v.px; g.in_ax = v.py * v.px;
        g.out_lane = 1; g.out_outer = v.px; g.out_ax = v.py * v.px;
    } else if (p == 1) {
        g.n_lane = v.px; g.n_outer = v.nz; g.n_ax = v.ny;
        g.in_outer = v.py * v.px; g.in_ax = v.px;
        g.out_lane = v.py; g.out_outer = v.px * v.py; g.out_ax = 1;
    } else {
        g.n_lane = v.py; g.n_outer = v.nz; g.n_ax = v.nx;
        g.in_outer = v.px * v.py; g.in_ax = v.py;
        g.out_lane = v.px; g.out_outer = v.py * v.px; g.out_ax = 1;
    }
    return g;
}

static void set_segments(LineGeom &g, int seg) {
    const int64_t full = ((g.n_ax + 15) / 16) * 16;
    g.seg = (seg <= 0 || seg >= full) ? full : seg;
    g.nseg = (g.n_ax + g.seg - 1) / g.seg;
}

// Build descriptors and work lists for segment length `seg` and upload them.
static int build_tables(mxb_batch_plan *P, int seg) {
    const int n = P->n;
    std::vector<BatchPassDesc> d(6 * (size_t)n);
    std::vector<BatchItem> lists[3];
    for (int v = 0; v < n; ++v) {
        const VolLayout &L = P->lay[v];
        for (int parity = 0; parity < 2; ++parity) {
            float *cur = P->st[parity], *tmp = P->st[1 - parity];
            cur += L.st_off;
            tmp += L.st_off;
            // z: cur -> tmp, y: tmp -> cur (T order), x: cur -> tmp + illuminate
            const float *ins[3] = {cur, tmp, cur};
            float *outs[3] = {tmp, cur, tmp};
            for (int p = 0; p < 3; ++p) {
                BatchPassDesc &D = d[(size_t)(parity * 3 + p) * n + v];
                D.in = ins[p];
                D.g = pitched_geom(L, p);
                set_segments(D.g, seg);
                const int64_t wpr = mask_words_per_row(L.px);
                D.e = p == 2 ? Epilogue{P->masks + L.mask_off, outs[p], nullptr,
                                        MaskBits{P->bits + L.bits_off, 1, wpr * L.py, L.py}}
                             : Epilogue{nullptr, outs[p], nullptr, MaskBits{}};
                D.bucket = L.bucket;
                D.taps_off = 0;
            }
        }
    }
    // taps table: per volume, centred in its bucket capacity
    std::vector<float> table;
    std::vector<int32_t> toff(n);
    for (int v = 0; v < n; ++v) {
        const int RT = kBucketRT[P->lay[v].bucket], R = P->lay[v].R;
        toff[v] = (int32_t)table.size();
        for (int i = 0; i < 2 * RT + 1; ++i) {
            const int r = i - RT;
            table.push_back((r >= -R && r <= R) ? P->host_taps[v][R + r] : 0.0f);
        }
    }
    for (auto &D : d) D.taps_off = 0;
    for (int k = 0; k < 6; ++k)
        for (int v = 0; v < n; ++v) d[(size_t)k * n + v].taps_off = toff[v];
    // 16-bit descriptors (same geometry and items) and split taps
    std::vector<BatchLowpDesc> ld;
    std::vector<uint16_t> ltable;
    if (P->lowp) {
        ld.resize(6 * (size_t)n);
        const bool bf = P->precision == MXB_BF16;
        for (int v = 0; v < n; ++v) {
            const VolLayout &L = P->lay[v];
            const int RT = kLowpBucketRT[L.lbucket];
            const int32_t off = (int32_t)ltable.size();
            ltable.resize(ltable.size() + 2 * (2 * RT + 1));
            split_taps_into(P->host_taps[v].data(), L.R, RT, bf, ltable.data() + off, ltable.data() + off + 2 * RT + 1);
            for (int parity = 0; parity < 2; ++parity) {
                uint16_t *cur = reinterpret_cast<uint16_t *>(P->st[parity] + L.st_off);
                uint16_t *tmp = reinterpret_cast<uint16_t *>(P->st[1 - parity] + L.st_off);
                const uint16_t *ins[3] = {cur, tmp, cur};
                uint16_t *outs[3] = {tmp, cur, tmp};
                for (int p = 0; p < 3; ++p) {
                    const BatchPassDesc &F = d[(size_t)(parity * 3 + p) * n + v];
                    BatchLowpDesc &D = ld[(size_t)(parity * 3 + p) * n + v];
                    D.in = ins[p];
                    D.out = outs[p];
                    D.mask = F.e.mask;
                    D.bits = F.e.bits;
                    D.g = F.g;
                    D.bucket = L.lbucket;
                    D.taps_off = off;
                }
            }
        }
    }
    for (int p = 0; p < 3; ++p) {
        std::vector<std::pair<int64_t, BatchItem>> tmp;
        for (int v = 0; v < n; ++v) {
            const LineGeom &g = d[(size_t)p * n + v].g;
            const int64_t nlb = (g.n_lane + 63) / 64;
            for (int64_t o = 0; o < g.n_outer; ++o)
                for (int64_t lb = 0; lb < nlb; ++lb)
                    for (int64_t s = 0; s < g.nseg; ++s) {
                        const int64_t a0 = s * g.seg;
                        const int64_t len = std::min(g.seg, g.n_ax - a0);
                        const int64_t cost = ((len + 15) / 16 + LdsShape<8>::W) * chunks_of(P->lay[v].bucket);
                        tmp.push_back({cost, BatchItem{v, (int32_t)s, (int32_t)lb, (int32_t)o}});
                    }
        }
        std::stable_sort(tmp.begin(), tmp.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
        lists[p].reserve(tmp.size());
        for (auto &t : tmp) lists[p].push_back(t.second);
    }
    const int total = (int)(lists[0].size() + lists[1].size() + lists[2].size());
    if (total > P->item_cap || table.size() > (size_t)(2 * 65 * n + 16) || ltable.size() > (size_t)(4 * 65 * n + 16)) {
        set_error("batch work list exceeds its reserved space");
        return EDEVICE;
    }
    int64_t off = 0;
    for (int p = 0; p < 3; ++p) {
        P->item_off[p] = off;
        P->n_items[p] = (int)lists[p].size();
        MXB_BHIP(hipMemcpyAsync(P->items + off, lists[p].data(), lists[p].size() * sizeof(BatchItem),
                                hipMemcpyHostToDevice, P->stream));
        off += (int64_t)lists[p].size();
    }
    MXB_BHIP(hipMemcpyAsync(P->descs, d.data(), d.size() * sizeof(BatchPassDesc), hipMemcpyHostToDevice, P->stream));
    MXB_BHIP(hipMemcpyAsync(P->taps, table.data(), table.size() * sizeof(float), hipMemcpyHostToDevice, P->stream));
    if (P->lowp) {
        MXB_BHIP(hipMemcpyAsync(P->ldescs, ld.data(), ld.size() * sizeof(BatchLowpDesc), hipMemcpyHostToDevice,
                                P->stream));
        MXB_BHIP(hipMemcpyAsync(P->ltaps, ltable.data(), ltable.size() * sizeof(uint16_t), hipMemcpyHostToDevice,
                                P->stream));
    }
    MXB_BHIP(hipStreamSynchronize(P->stream));
    return OK;
}

static hipError_t launch_pass(mxb_batch_plan *P, int parity, int p, int grid) {
    const int n_items = P->n_items[p];
    if (n_items == 0) return hipSuccess;
    const unsigned g = (unsigned)std::min(grid, n_items);
    const BatchPassDesc *d = P->descs + (size_t)(parity * 3 + p) * P->n;
    const BatchItem *it = P->items + P->item_off[p];
    if (p == 0)
        hipLaunchKernelGGL(k_batched_pass<0>, dim3(g), dim3(256), P->smem_bytes[0], P->stream, d, it, n_items, P->taps);
    else if (p == 1)
        hipLaunchKernelGGL(k_batched_pass<1>, dim3(g), dim3(256), P->smem_bytes[1], P->stream, d, it, n_items, P->taps);
    else
        hipLaunchKernelGGL(k_batched_pass<2>, dim3(g), dim3(256), P->smem_bytes[2], P->stream, d, it, n_items, P->taps);
    return hipGetLastError();
}

static hipError_t launch_lowp_pass(mxb_batch_plan *P, int parity, int p, int grid) {
    const int n_items = P->n_items[p];
    if (n_items == 0) return hipSuccess;
    const unsigned g = (unsigned)std::min(grid, n_items);
    const BatchLowpDesc *d = P->ldescs + (size_t)(parity * 3 + p) * P->n;
    const BatchItem *it = P->items + P->item_off[p];
    const size_t sm = (size_t)P->lsmem_bytes[p];
    const bool bf = P->precision == MXB_BF16;
#define MXB_LOWP_LAUNCH(F, B) \
    hipLaunchKernelGGL((k_batched_lowp_pass<F, B>), dim3(g), dim3(256), sm, P->stream, d, it, n_items, P->ltaps)
    if (p == 0) {
        if (bf) MXB_LOWP_LAUNCH(0, true); else MXB_LOWP_LAUNCH(0, false);
    } else if (p == 1) {
        if (bf) MXB_LOWP_LAUNCH(1, true); else MXB_LOWP_LAUNCH(1, false);
    } else {
        if (bf) MXB_LOWP_LAUNCH(2, true); else MXB_LOWP_LAUNCH(2, false);
    }
#undef MXB_LOWP_LAUNCH
    return hipGetLastError();
}

// Enqueue the whole persistent-form diffusion on P->stream.
static int enqueue_persistent(mxb_batch_plan *P) {
    if (P->lowp && P->reps > 1) {
        // repetitions 1 .. reps-1 in 16-bit state (z: s0 -> s1, y: s1 -> s0,
        // x: s0 -> s1, swap), widen, then one FP32 repetition
        const bool bf = P->precision == MXB_BF16;
        for (int v = 0; v < P->n; ++v) {
            const VolLayout &L = P->lay[v];
            MXB_BHIP(launch_init_lowp(P->masks + L.mask_off, reinterpret_cast<uint16_t *>(P->st[0] + L.st_off),
                                      L.nz * L.py * L.px, bf, P->stream));
        }
        int parity = 0;
        for (int64_t t = 0; t < P->reps - 1; ++t) {
            for (int p = 0; p < 3; ++p) MXB_BHIP(launch_lowp_pass(P, parity, p, P->tune.grid));
            parity ^= 1;
        }
        for (int v = 0; v < P->n; ++v) {
            const VolLayout &L = P->lay[v];
            MXB_BHIP(launch_lowp_to_f32(reinterpret_cast<const uint16_t *>(P->st[parity] + L.st_off),
                                        P->st[1 - parity] + L.st_off, L.nz * L.py * L.px, bf, P->stream));
        }
        parity ^= 1;
        for (int p = 0; p < 3; ++p) MXB_BHIP(launch_pass(P, parity, p, P->tune.grid));
        parity ^= 1;
        for (int v = 0; v < P->n; ++v) {
            const VolLayout &L = P->lay[v];
            const mxb_device_volume &V = P->vols[v];
            const int64_t n = L.nz * L.ny * L.nx;
            hipLaunchKernelGGL(k_unpad_out, dim3(blocks_for(n)), dim3(256), 0, P->stream, P->st[parity] + L.st_off,
                               L.nz, L.ny, L.nx, L.py, L.px, V.dev_out_u16, V.dev_out_f32);
            MXB_BHIP(hipGetLastError());
        }
        return OK;
    }
    for (int v = 0; v < P->n; ++v) {
        const VolLayout &L = P->lay[v];
        const int64_t n = L.nz * L.py * L.px;
        hipLaunchKernelGGL(k_init_pitched, dim3(blocks_for(n)), dim3(256), 0, P->stream, P->masks + L.mask_off,
                           P->st[0] + L.st_off, n);
        MXB_BHIP(hipGetLastError());
    }
    int parity = 0;
    for (int64_t t = 0; t < P->reps; ++t) {
        for (int p = 0; p < 3; ++p) MXB_BHIP(launch_pass(P, parity, p, P->tune.grid));
        parity ^= 1;
    }
    for (int v = 0; v < P->n; ++v) {
        const VolLayout &L = P->lay[v];
        const mxb_device_volume &V = P->vols[v];
        const int64_t n = L.nz * L.ny * L.nx;
        hipLaunchKernelGGL(k_unpad_out, dim3(blocks_for(n)), dim3(256), 0, P->stream, P->st[parity] + L.st_off,
                           L.nz, L.ny, L.nx, L.py, L.px, V.dev_out_u16, V.dev_out_f32);
        MXB_BHIP(hipGetLastError());
    }
    return OK;
}

// Time one repetition (three passes) with the current tables.
static int time_one_rep(mxb_batch_plan *P, int grid, double *ms) {
    hipEvent_t a, b;
    MXB_BHIP(hipEventCreate(&a));
    MXB_BHIP(hipEventCreate(&b));
    // 16-bit plans spend all but one repetition in the 16-bit passes
    auto pass = [&](int p) { return P->lowp ? launch_lowp_pass(P, 0, p, grid) : launch_pass(P, 0, p, grid); };
    for (int p = 0; p < 3; ++p) MXB_BHIP(pass(p));  // warm
    MXB_BHIP(hipEventRecord(a, P->stream));
    for (int p = 0; p < 3; ++p) MXB_BHIP(pass(p));
    MXB_BHIP(hipEventRecord(b, P->stream));
    MXB_BHIP(hipEventSynchronize(b));
    float f = 0;
    MXB_BHIP(hipEventElapsedTime(&f, a, b));
    *ms = f;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return OK;
}

static int autotune(mxb_batch_plan *P, int cus) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    load_tune_file();
    auto it = g_tune.find(P->key);
    if (it != g_tune.end()) {
        P->tune = it->second;
        return build_tables(P, P->tune.seg);
    }
    // init state so that the timed repetitions read finite values
    for (int v = 0; v < P->n; ++v) {
        const VolLayout &L = P->lay[v];
        const int64_t n = L.nz * L.py * L.px;
        if (P->lowp) {
            MXB_BHIP(launch_init_lowp(P->masks + L.mask_off, reinterpret_cast<uint16_t *>(P->st[0] + L.st_off), n,
                                      P->precision == MXB_BF16, P->stream));
        } else {
            hipLaunchKernelGGL(k_init_pitched, dim3(blocks_for(n)), dim3(256), 0, P->stream,
                               P->masks + L.mask_off, P->st[0] + L.st_off, n);
            MXB_BHIP(hipGetLastError());
        }
    }
    TuneChoice best;
    best.ms = 1e30;
    const int segs[] = {0, 256, 128};
    const int per_cu[] = {2, 3, 4, 6, 8};
    for (int seg : segs) {
        int rc = build_tables(P, seg);
        if (rc != OK) return rc;
        for (int k : per_cu) {
            const int grid = k * cus;
            // the fastest of three timed repetitions (one is noisy)
            double ms = 1e30;
            for (int trial = 0; trial < 3; ++trial) {
                double t = 0;
                rc = time_one_rep(P, grid, &t);
                if (rc != OK) return rc;
                ms = std::min(ms, t);
            }
            if (ms < best.ms) {
                best.ms = ms;
                best.grid = grid;
                best.seg = seg;
            }
        }
    }
    P->tune = best;
    g_tune[P->key] = best;
    save_tune_file();
    return build_tables(P, best.seg);
}

static int capture_graph(mxb_batch_plan *P) {
    MXB_BHIP(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_persistent(P);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(P->stream, &g);
    if (rc != OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    MXB_BHIP(e);
    P->graph = g;
    MXB_BHIP(hipGraphInstantiate(&P->exec, g, nullptr, nullptr, 0));
    return OK;
}

static int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

}  // namespace mxb

using namespace mxb;

extern "C" {

int mxb_batch_create(const mxb_device_volume *vols, int n, int64_t repetitions, int precision,
                     void *stream, mxb_batch_plan **plan_out) {
    if (plan_out == nullptr || n < 0 || (n > 0 && vols == nullptr) || repetitions < 0) {
        set_error("mxb_batch_create: invalid arguments");
        return EINVAL_;
    }
    if (precision < MXB_FP32 || precision > MXB_FP32_SPLIT) {
        set_error("unknown precision");
        return EINVAL_;
    }
    *plan_out = nullptr;
    auto *P = new mxb_batch_plan();
    P->n = n;
    P->reps = repetitions;
    P->precision = precision;
    P->stream = (hipStream_t)stream;
    if (P->stream == nullptr) {
        // graph capture needs a stream of its own
        if (hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking) != hipSuccess) {
            set_error("mxb_batch_create: stream creation failed");
            delete P;
            return EDEVICE;
        }
        P->own_stream = true;
    }
    P->vols.assign(vols, vols + n);
    // the persistent form serves FP32 (MFMA) and BF16 / FP16 volumes with R <= 64
    P->lowp = precision == MXB_BF16 || precision == MXB_FP16;
    P->persistent = precision == MXB_FP32 || P->lowp;
    int64_t st_total = 0, mask_total = 0, bits_total = 0;
    std::string key = "b" + std::to_string(n) + "p" + std::to_string(precision);
    for (int v = 0; v < n; ++v) {
        const mxb_device_volume &V = vols[v];
        if (V.nz < 1 || V.ny < 1 || V.nx < 1 || V.kernel == nullptr || V.kernel_size < 1 ||
            V.kernel_size % 2 == 0 || V.dev_mask == nullptr || (V.dev_out_u16 == nullptr && V.dev_out_f32 == nullptr)) {
            set_error("volume " + std::to_string(v) + ": invalid descriptor");
            mxb_batch_destroy(P);
            return EINVAL_;
        }
        VolLayout L{};
        L.nz = V.nz; L.ny = V.ny; L.nx = V.nx;
        L.px = round_up(V.nx, pitch_elements(precision));
        L.py = round_up(V.ny, pitch_elements(precision));
        L.R = (int)(V.kernel_size / 2);
        if (L.R > 64) P->persistent = false;
        L.bucket = kNumBuckets - 1;
        for (int b = 0; b < kNumBuckets; ++b)
            if (L.R <= kBucketRT[b]) { L.bucket = b; break; }
        L.lbucket = kNumLowpBuckets - 1;
        for (int b = 0; b < kNumLowpBuckets; ++b)
            if (L.R <= kLowpBucketRT[b]) { L.lbucket = b; break; }
        L.st_off = st_total;
        // the per-volume form runs mxb_device_diffusion on these buffers
        const int64_t sb = mxb_state_bytes(precision, V.nz, V.ny, V.nx);
        st_total += round_up(std::max(V.nz * L.py * L.px, (sb + 3) / 4), 64);
        L.mask_off = mask_total;
        mask_total += round_up(V.nz * L.py * L.px, 256);
        L.bits_off = bits_total;
        bits_total += round_up(mask_bits_bytes(V.nz, L.py, L.px) / 4, 64);
        P->lay.push_back(L);
        P->host_taps.emplace_back(V.kernel, V.kernel + V.kernel_size);
        key += "_" + std::to_string(V.nz) + "x" + std::to_string(V.ny) + "x" + std::to_string(V.nx) + "r" +
               std::to_string(L.R);
    }
    P->key = key;
    if (n == 0) {
        *plan_out = P;
        return OK;
    }
    // workspace: masks | state0 | state1 | taps | descriptors | items
    int64_t items_cap = 0;
    for (auto &L : P->lay) {
        // an upper bound on items of any pass for any segment choice (seg >= 128)
        const int64_t m = std::max({L.nz, L.ny, L.nx});
        items_cap += 3 * ((std::max(L.px, L.py) + 63) / 64) * std::max(L.ny, L.nz) * ((m + 127) / 128);
    }
    P->item_cap = (int)std::min<int64_t>(items_cap, INT32_MAX / 2);
    const int64_t taps_cap = 2 * 65 * (int64_t)n + 16;
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    const size_t o_mask = carve((size_t)mask_total);
    const size_t o_bits = carve((size_t)bits_total * 4);
    const size_t o_s0 = carve((size_t)st_total * 4);
    const size_t o_s1 = carve((size_t)st_total * 4);
    const size_t o_taps = carve((size_t)taps_cap * 4);
    const size_t o_desc = carve(6 * (size_t)n * sizeof(BatchPassDesc));
    const size_t o_ldesc = carve(P->lowp ? 6 * (size_t)n * sizeof(BatchLowpDesc) : 0);
    const size_t o_ltaps = carve(P->lowp ? (size_t)(4 * 65 * (int64_t)n + 16) * sizeof(uint16_t) : 0);
    const size_t o_items = carve((size_t)P->item_cap * sizeof(BatchItem));
    hipError_t e = hipMalloc(&P->ws, off);
    if (e != hipSuccess) {
        set_error(std::string("batch workspace: ") + hipGetErrorString(e));
        mxb_batch_destroy(P);
        return e == hipErrorOutOfMemory ? ENOMEM_ : EDEVICE;
    }
    char *base = (char *)P->ws;
    P->masks = (uint8_t *)(base + o_mask);
    P->bits = (uint32_t *)(base + o_bits);
    P->st[0] = (float *)(base + o_s0);
    P->st[1] = (float *)(base + o_s1);
    P->taps = (float *)(base + o_taps);
    P->descs = (BatchPassDesc *)(base + o_desc);
    if (P->lowp) {
        P->ldescs = (BatchLowpDesc *)(base + o_ldesc);
        P->ltaps = (uint16_t *)(base + o_ltaps);
    }
    P->items = (BatchItem *)(base + o_items);
    int rc = OK;
    auto fail = [&](int code) {
        mxb_batch_destroy(P);
        return code;
    };
    // pitched masks (pad columns zero): set-up work, outside any graph
    if ((e = hipMemsetAsync(P->masks, 0, (size_t)mask_total, P->stream)) != hipSuccess) {
        set_error(std::string("hipMemsetAsync: ") + hipGetErrorString(e));
        return fail(EDEVICE);
    }
    for (int v = 0; v < n; ++v) {
        const VolLayout &L = P->lay[v];
        hipLaunchKernelGGL(k_pitch_mask, dim3(blocks_for(L.nz * L.ny * L.nx)), dim3(256), 0, P->stream,
                           vols[v].dev_mask, L.nz, L.ny, L.nx, L.py, L.px, P->masks + L.mask_off);
        if ((e = hipGetLastError()) != hipSuccess) {
            set_error(std::string("k_pitch_mask: ") + hipGetErrorString(e));
            return fail(EDEVICE);
        }
        if ((e = launch_pack_mask_bits(P->masks + L.mask_off, L.nz, L.py, L.px, L.px, L.py * L.px,
                                       P->bits + L.bits_off, P->stream)) != hipSuccess) {
            set_error(std::string("k_pack_mask_bits: ") + hipGetErrorString(e));
            return fail(EDEVICE);
        }
    }
    if (P->persistent) {
        for (auto &L : P->lay) {
            P->smem_bytes[0] = std::max(P->smem_bytes[0], smem_bytes_of<0>(L.bucket));
            P->smem_bytes[1] = std::max(P->smem_bytes[1], smem_bytes_of<1>(L.bucket));
            P->smem_bytes[2] = std::max(P->smem_bytes[2], smem_bytes_of<2>(L.bucket));
            if (P->lowp) {
                P->lsmem_bytes[0] = std::max(P->lsmem_bytes[0], lowp_smem_bytes_of<0>(L.lbucket));
                P->lsmem_bytes[1] = std::max(P->lsmem_bytes[1], lowp_smem_bytes_of<1>(L.lbucket));
                P->lsmem_bytes[2] = std::max(P->lsmem_bytes[2], lowp_smem_bytes_of<2>(L.lbucket));
            }
        }
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess) {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
        }
        if ((rc = autotune(P, cus)) != OK) return fail(rc);
        if ((rc = capture_graph(P)) != OK) return fail(rc);
    }
    *plan_out = P;
    return OK;
}

int mxb_batch_run(mxb_batch_plan *P, int mode) {
    if (P == nullptr) {
        set_error("null plan");
        return EINVAL_;
    }
    if (P->n == 0) return OK;
    if (mode == 0 && P->persistent) {
        MXB_BHIP(hipGraphLaunch(P->exec, P->stream));
        return OK;
    }
    // per-volume launches (also the form used for the other precisions), with
    // the persistent passes' three-pass structure
    for (int v = 0; v < P->n; ++v) {
        const mxb_device_volume &V = P->vols[v];
        const VolLayout &L = P->lay[v];
        int rc = device_diffusion(V.dev_mask, V.nz, V.ny, V.nx, V.kernel, V.kernel_size, P->reps,
                                  P->precision, P->st[0] + L.st_off, P->st[1] + L.st_off, V.dev_out_u16,
                                  V.dev_out_f32, P->stream, false);
        if (rc != OK) return rc;
    }
    return OK;
}

int mxb_batch_info(const mxb_batch_plan *P, int *persistent, int *grid, int *seg, double *tuned_ms) {
    if (P == nullptr) {
        set_error("null plan");
        return EINVAL_;
    }
    if (persistent) *persistent = P->persistent ? 1 : 0;
    if (grid) *grid = P->tune.grid;
    if (seg) *seg = P->tune.seg;
    if (tuned_ms) *tuned_ms = P->tune.ms;
    return OK;
}

int mxb_batch_sync(mxb_batch_plan *P) {
    if (P == nullptr) {
        set_error("null plan");
        return EINVAL_;
    }
    MXB_BHIP(hipStreamSynchronize(P->stream));
    return OK;
}

int mxb_batch_destroy(mxb_batch_plan *P) {
    if (P == nullptr) return OK;
    if (P->stream) (void)hipStreamSynchronize(P->stream);
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    if (P->ws) (void)hipFree(P->ws);
    if (P->own_stream) (void)hipStreamDestroy(P->stream);
    delete P;
    return OK;
}

// Host-buffer batch: upload, run the plan (persistent graph for FP32), download.
int mxb_diffusion_batched(const mxb_volume_desc *volumes, int n, int64_t repetitions, int precision,
                          int verbose) {
    (void)verbose;
    if (n < 0 || (n > 0 && volumes == nullptr)) {
        set_error("volumes must hold n descriptors");
        return EINVAL_;
    }
    if (n == 0) return OK;
    std::vector<mxb_device_volume> dv(n);
    std::vector<void *> bufs;
    auto cleanup = [&]() {
        for (void *b : bufs) (void)hipFree(b);
    };
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        const mxb_volume_desc &v = volumes[i];
        if (v.nz < 1 || v.ny < 1 || v.nx < 1 || v.mask == nullptr ||
            (v.output_u16 == nullptr && v.output_f32 == nullptr)) {
            set_error("volume " + std::to_string(i) + ": invalid descriptor");
            cleanup();
            return EINVAL_;
        }
        const int64_t nv = v.nz * v.ny * v.nx;
        total += nv;
        void *m = nullptr, *q = nullptr, *f = nullptr;
        hipError_t e = hipMalloc(&m, nv);
        if (e == hipSuccess && v.output_u16) e = hipMalloc(&q, nv * 2);
        if (e == hipSuccess && v.output_f32) e = hipMalloc(&f, nv * 4);
        if (m) bufs.push_back(m);
        if (q) bufs.push_back(q);
        if (f) bufs.push_back(f);
        if (e != hipSuccess) {
            set_error(std::string("batch buffers: ") + hipGetErrorString(e));
            cleanup();
            return e == hipErrorOutOfMemory ? ENOMEM_ : EDEVICE;
        }
        if ((e = hipMemcpy(m, v.mask, nv, hipMemcpyHostToDevice)) != hipSuccess) {
            set_error(std::string("hipMemcpy: ") + hipGetErrorString(e));
            cleanup();
            return EDEVICE;
        }
        dv[i] = mxb_device_volume{(const uint8_t *)m, v.nz, v.ny, v.nx, v.kernel, v.kernel_size,
                                  (uint16_t *)q, (float *)f};
    }
    mxb_batch_plan *P = nullptr;
    int rc = mxb_batch_create(dv.data(), n, repetitions, precision, nullptr, &P);
    if (rc == OK) rc = mxb_batch_run(P, 0);
    if (rc == OK) rc = mxb_batch_sync(P);
    if (rc == OK) {
        for (int i = 0; i < n && rc == OK; ++i) {
            const mxb_volume_desc &v = volumes[i];
            const int64_t nv = v.nz * v.ny * v.nx;
            hipError_t e = hipSuccess;
            if (v.output_u16) e = hipMemcpy(v.output_u16, dv[i].dev_out_u16, nv * 2, hipMemcpyDeviceToHost);
            if (e == hipSuccess && v.output_f32) e = hipMemcpy(v.output_f32, dv[i].dev_out_f32, nv * 4, hipMemcpyDeviceToHost);
            if (e != hipSuccess) {
                set_error(std::string("hipMemcpy: ") + hipGetErrorString(e));
                rc = EDEVICE;
            }
        }
    }
    mxb_batch_destroy(P);
    cleanup();
    (void)total;
    return rc;
}

}  // extern "C"
