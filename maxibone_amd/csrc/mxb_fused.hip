// Fused x -> y pass of the masked 3D diffusion (gfx950), 16-bit MFMA forms.
//
// A repetition of the 16-bit and split forms runs as two kernels: the z-pass
// (the direct-load 16-bit body, A -> B) and this one (B -> A), which applies
// the x-pass and then the y-pass to one z-plane strip with the intermediate in
// LDS, so the (z, x, y) round trip of the three-pass form disappears.  The
// separable passes commute, so z, x, y computes the same correlation as the
// reference's z, y, x (src/lib/cpp/cpu/diffusion.cc:314-320), with its own
// rounding; illumination (cpu:295-302) follows the last pass, as there.
//
// Block = 512 threads (8 waves) on plane z, output columns [x0, x0 + TX) and
// output rows [ya, yb).  It marches along y in steps of 16 rows:
//   1. direct loads bring input rows ya - RT + 16 j .. +15, columns
//      [x0 - HALO, x0 + TX + HALO), into a staging buffer (D + 1 buffers:
//      steps j + 1 .. j + D load while step j computes); pieces outside the
//      halo the radius needs, and everything outside the volume, read a zero
//      block;
//   2. the x-pass: 16x16x32 MFMA tiles with lines = rows and positions = x,
//      operands read straight from the staged rows (no transpose), results
//      rounded to the state format and written into half j % 2 of slot
//      j / 2 of a y-ring of W slots laid out as the 16-bit body's slots (32
//      positions x 64 lines per 4 KB, one per 64-column sub-ring);
//   3. after every second step, the y-pass of the next 32-row output pair
//      from W slots with the 16-bit body's transposed operand reads (lines =
//      x, positions = y);
//   4. the epilogue stages the pair's 32 rows x TX outputs in LDS, sets mask
//      voxels to 1 from the mask bits (direct-loaded every second step), and
//      writes whole 16-byte row pieces.
// Arithmetic, tap split and tile chunking are the 16-bit body's (mxb_lowp.h);
// the split form (FP16 hi + lo planes, SplitIO) runs three MFMAs per chunk.
//
// Wait counts: the loads of step j + D are issued after the barrier that
// follows the wait for step j, and the only vector-memory operations are the
// steps' loads (GL per thread) and the epilogues' stores (ST per thread, all
// unconditional: stores outside the volume go to g_dma_sink), so the wait
// before step j leaves outstanding the loads of steps j + 1 .. j + D - 1 and
// the stores of the epilogues that ran after step j's loads were issued.
#include "maxibone_diffusion.h"
#include "mxb_dma.h"
#include "mxb_lowp.h"
#include "mxb_lowp_dma_body.h"

namespace mxb {

struct XYGeom {
    int64_t nz, ny, nx;
    int64_t seg, nseg;   // output rows per block, blocks along y
    int64_t nstrip;      // column strips of TX
    int64_t wpr;         // mask words per row
};

constexpr int kXYWaves = 8;  // waves per block

// Staged rows are XOR-swizzled in 16-byte pieces over groups of this many
// consecutive rows (a power of two), so that the x-pass's 16-byte operand
// reads of that many consecutive rows (one per lane) fall in distinct banks.
#ifndef MXB_XY_SWZ
#define MXB_XY_SWZ 16
#endif
constexpr int kXYSwz = MXB_XY_SWZ - 1;

// Epilogue form.  Wave epilogue: each wave stages and stores its own
// 16-column x 32-row y-pass tile through a wave-private LDS area, with no
// block barrier before the stores (rows leave in 32-byte pieces per wave).
// Block epilogue: the block stages the pair's 32 rows x TX outputs and
// stores whole rows after a barrier.  1 (measured on 1024^3, R=40): the wave
// epilogue for the split form (fused pass 2.52 against 2.64 ms), the block
// epilogue for the 16-bit forms (1.375 against 1.42 ms); 0: block always;
// 2: wave always.
#ifndef MXB_XY_WAVE_EPI
#define MXB_XY_WAVE_EPI 1
#endif

// Per-form shape of the 16-bit and split kernels (measured on 1024^3, R=40,
// ms per fused pass):
// * split form: its own epilogue stage, 3 steps of prefetch, one block per
//   CU (158 KB of LDS), operand reads issued before the MFMAs (2.43 against
//   2.52 interleaved);
// * 16-bit forms: the epilogue stages in the staging buffer its own step has
//   just consumed (the x-pass read it before the y-pass barrier; the next
//   step's loads refill it only after that step's first barrier), 1 step of
//   prefetch, so the block needs 51 KB and three blocks share a CU (six waves
//   per SIMD, at most 85 VGPRs; the early operand reads would spill there):
//   1.26 against 1.35 with 3 steps at two blocks per CU.
// MXB_XY_PRELOAD forces the operand-read order (1 early, 0 interleaved) for
// both forms.
#ifdef MXB_XY_PRELOAD
#define MXB_XY_PRELOAD_FOR(split) (MXB_XY_PRELOAD)
#else
#define MXB_XY_PRELOAD_FOR(split) (split)
#endif
static_assert((MXB_XY_SWZ & kXYSwz) == 0 && MXB_XY_SWZ <= 16, "swizzle group: a power of two up to 16");

template <int RT, int TX, int D, bool SPLIT, int FIN>
struct XYShape {
    static constexpr int NW = kXYWaves;
    static constexpr int P = SPLIT ? 2 : 1;           // state planes
    static constexpr int C = LowpChunks<RT>::C;       // 32-position chunks per tile
    static constexpr int W = C + 1;                   // ring slots a tile pair reads
    static constexpr int HALO = 64;                   // staged halves each side of the strip
    static constexpr int ROWH = TX + 2 * HALO;        // staged halves per row
    static constexpr int PR = ROWH / 8;               // 16-byte pieces per staged row
    static constexpr int SB = 16 * ROWH * 2;          // staging bytes per plane and step (16 rows)
    static constexpr int NI = 16 * PR / (64 * NW);    // direct loads per wave, plane and step
    static constexpr int NSUB = TX / 64;              // 64-column sub-rings
    static constexpr int SLOT = NSUB * kSlotBytes;    // ring bytes per plane and 32-row slot
    static constexpr int MW = TX / 32;                // mask words per row of the strip
    static constexpr int MHALF = MW * 64 > 256 ? MW * 64 : 256;  // mask bytes per 16 rows
    static constexpr int MLOADS = (MW + 3) / 4;       // waves loading mask words per step
    static constexpr int NT = TX / (16 * NW);         // x-pass column tiles per wave
    static constexpr int NQ = TX / (16 * NW);         // y-pass 16-column groups per wave
    static constexpr int MSLOT = 2 * MHALF;           // mask bytes per pair
    static constexpr int PLO = (HALO - (RT + 7) / 8 * 8) / 8;  // first piece the radius needs
    static constexpr int STAGE = (D + 1) * P * SB;    // D steps in flight while one computes
    static constexpr int RING = W * P * SLOT;
    // the block epilogue's stage: the consumed staging buffer (16-bit forms;
    // unpadded rows, 16-byte pieces swizzled by row) or an area of its own
    static constexpr bool ALIAS = !SPLIT;
    static constexpr bool PRELOAD = MXB_XY_PRELOAD_FOR(SPLIT);
    static constexpr int SPITCH16 = 2 * TX + (ALIAS ? 0 : 16);  // block stage pitch (bytes), 16-bit
    static constexpr int SPITCHF = 4 * TX + (ALIAS ? 0 : 16);   // block stage pitch (bytes), FP32
    static constexpr int WP16 = 48;                   // wave stage pitch (bytes), 16-bit: 16 columns + pad
    static constexpr int WPF = 80;                    // wave stage pitch (bytes), FP32
    static constexpr int WEST = 32 * (FIN != 0 ? WPF : P * WP16);  // wave stage bytes
    static constexpr bool WEPI = MXB_XY_WAVE_EPI == 2 || (MXB_XY_WAVE_EPI == 1 && SPLIT);
    static constexpr int EST_NEED = WEPI ? NW * WEST : 32 * (FIN != 0 ? SPITCHF : P * SPITCH16);
    static constexpr int EST = ALIAS && !WEPI ? 0 : EST_NEED;
    static_assert(!(ALIAS && !WEPI) || (EST_NEED <= P * SB && (2 * TX / 16) % 16 == 0),
                  "the block stage fits one step's staging buffer in whole 16-piece swizzle groups");
    static constexpr int MRING = (D + 1) * MSLOT + 256;  // + the target of the idle waves' mask loads
    static constexpr int TAPS = 2 * LowpChunks<RT>::TW * 2;
    static constexpr int BYTES = STAGE + RING + EST + MRING + TAPS;
    static constexpr int GL = P * NI + 1;             // direct loads per wave and step
    // stores per thread and pair: 32 rows x TX outputs in 16-byte pieces
    static constexpr int PIECES16 = 32 * TX / 8 / (64 * NW);  // 16-bit pieces per thread
    static constexpr int ST = FIN != 0 ? ((FIN & 1) ? 2 * PIECES16 : 0) + ((FIN & 2) ? PIECES16 : 0)
                                       : P * PIECES16;
    static_assert(TX % (16 * NW) == 0 && (16 * PR) % (64 * NW) == 0 && PIECES16 >= 1,
                  "whole tiles, loads and stores per wave");
    static_assert(PR % MXB_XY_SWZ == 0, "the swizzle keeps a piece within its row");
    static_assert(!WEPI || (NQ == 1 && PIECES16 == 1),
                  "wave epilogue: one 16-column group per wave, one 8-column piece per lane");
    static_assert(RT % 4 == 0, "radius bucket");
    static_assert(FIN == 0 || SPLIT, "final outputs are the split form's");
    static_assert(MLOADS <= 4, "mask words of 16 rows in one load per wave");
    // the last tile's last chunk ends at staged half HALO - RT + TX + 32 C - 16;
    // where that passes the row end (radius bucket 64, whose 32-position chunks
    // overhang the 16 + 2 RT positions a tile needs by 16), the x-pass reads
    // zeros for the overhang, which meets only zero taps.  The positions a
    // tile needs end at HALO + TX + RT, inside the row when RT <= HALO.
    static constexpr bool GUARD = HALO - RT + TX + 32 * C - 16 > ROWH;
    static constexpr bool READS_IN_ROW = RT <= HALO;
    static constexpr int NST = D / 2 + 1;  // most epilogues between a group and its wait (every second step)
    static_assert(D >= 1 && (D - 1) * GL + NST * ST < 64, "vmcnt range");
    static constexpr bool FITS = READS_IN_ROW && BYTES <= 160 * 1024;  // gfx950 LDS per workgroup
};

// x-pass products of operand read q: it is chunk c = q - i of tile t0 + 2 i
// (taps in chunk order per tile).
template <class L, int NT, int C, bool SPLIT, int NR>
__device__ __forceinline__ void xy_x_products(int q, const v8s (&xs)[NR], const v8s (&xls)[NR], const v8s (&ahi)[C],
                                              const v8s (&alo)[C], v4f (&acc)[NT]) {
#pragma unroll
    for (int i = 0; i < NT; ++i) {
        const int c = q - i;
        if (c < 0 || c >= C) continue;
        acc[i] = L::mfma(ahi[c], xs[q], acc[i]);
        acc[i] = L::mfma(alo[c], xs[q], acc[i]);
        if constexpr (SPLIT) acc[i] = L::mfma(ahi[c], xls[q], acc[i]);
    }
}

// Byte offset of byte b of row r in the block epilogue's stage of
// `pitch`-byte rows: padded rows of its own, or (aliased into a staging
// buffer) unpadded rows whose 16-byte pieces are XOR-swizzled by the row, so
// that the y-pass lanes' writes 4 rows apart fall in distinct banks.
template <bool ALIAS>
__device__ __forceinline__ int stage_off(int r, int b, int pitch) {
    if constexpr (ALIAS) return r * pitch + ((((b >> 4) ^ (r & 15)) << 4) | (b & 15));
    else return r * pitch + b;
}

// LDS ordering between the lanes of one wave: its LDS operations execute in
// order; the fences keep the compiler from moving them across.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Final outputs of one 8-column piece: mask voxels (bits of m) become 1, then
// the FP32 state (FIN & 1) and / or the uint16 result (FIN & 2) are stored;
// pieces outside the volume go to the sink (the store count stays fixed).
template <int FIN>
__device__ __forceinline__ void xy_store_fin(v4f (&v)[2], uint32_t m, bool ok, int64_t off, const SplitIO &io,
                                             float *sinkf, uint16_t *sink) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        if ((m >> e) & 1u) v[0][e] = 1.0f;
        if ((m >> (4 + e)) & 1u) v[1][e] = 1.0f;
    }
    if constexpr ((FIN & 1) != 0) {
        float *dst = ok ? io.fin_f32 + off : sinkf;
        store16<false>(dst, v[0]);
        store16<false>(dst + (ok ? 4 : 0), v[1]);
    }
    if constexpr ((FIN & 2) != 0) {
        uint32_t wq[4];
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            wq[2 * hh] = (uint32_t)quantise_u16_s(v[hh][0]) | ((uint32_t)quantise_u16_s(v[hh][1]) << 16);
            wq[2 * hh + 1] = (uint32_t)quantise_u16_s(v[hh][2]) | ((uint32_t)quantise_u16_s(v[hh][3]) << 16);
        }
        store16<false>(ok ? io.fin_u16 + off : sink, make_uint4(wq[0], wq[1], wq[2], wq[3]));
    }
}

// 16-bit state of one 8-column piece (v; the split form's low plane vl):
// halves whose mask bit is set become 1.0 (low plane 0), then stored.
template <bool BF, bool SPLIT>
__device__ __forceinline__ void xy_store_state(uint4 v, uint4 vl, uint32_t m, bool ok, int64_t off, uint16_t *out,
                                               const SplitIO &io, uint16_t *sink) {
    using L = Lowp<BF>;
    const uint32_t one2 = L::from_f32(1.0f) * 0x10001u;  // 1.0 in both halves
    uint32_t w[4] = {v.x, v.y, v.z, v.w};
    uint32_t wl[4] = {vl.x, vl.y, vl.z, vl.w};
#pragma unroll
    for (int e2 = 0; e2 < 4; ++e2) {
        const uint32_t sel = ((m >> (2 * e2)) & 1u) * 0xffffu | ((m >> (2 * e2 + 1)) & 1u) * 0xffff0000u;
        w[e2] = (w[e2] & ~sel) | (one2 & sel);
        wl[e2] &= ~sel;
    }
    store16<false>(ok ? out + off : sink, make_uint4(w[0], w[1], w[2], w[3]));
    if constexpr (SPLIT) store16<false>(ok ? out + io.out_lo + off : sink, make_uint4(wl[0], wl[1], wl[2], wl[3]));
}

// Steps of 16 rows: step j stages input rows ya - RT + 16 j .. +15 and the
// x-pass writes them to half j % 2 of ring slot (j / 2) % W; the y-pass of
// pair p (output rows ya + 32 p ..) runs after step 2 p + 2 W - 1, the step
// that completes slot p + W - 1.  Its mask words come with steps 2 p + 2 W - 2
// (rows 0 .. 15) and 2 p + 2 W - 1 (rows 16 .. 31).
template <int RT, int TX, int D, bool BF, bool SPLIT, int FIN>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(SPLIT ? 1 : 6))) void k_xy_fused(const uint16_t *__restrict__ in, uint16_t *__restrict__ out,
                                                  const uint32_t *__restrict__ bits, XYGeom g,
                                                  TapsL<2 * RT + 1> taps, SplitIO io) {
    using S = XYShape<RT, TX, D, SPLIT, FIN>;
    using L = Lowp<BF>;
    constexpr int P = S::P, C = S::C, W = S::W, PR = S::PR, SB = S::SB, NI = S::NI;
    constexpr int SLOT = S::SLOT, MW = S::MW, MSLOT = S::MSLOT, MHALF = S::MHALF, ST = S::ST, GL = S::GL;
    constexpr int TW = LowpChunks<RT>::TW;
    static_assert(!SPLIT || !BF, "the split form runs on FP16 MFMAs");
    static_assert(S::FITS, "shape fits the LDS");
    __shared__ __attribute__((aligned(16))) unsigned char smem[S::BYTES];
    unsigned char *const staging = smem;
    unsigned char *const ring = smem + S::STAGE;
    unsigned char *const est_own = ring + S::RING;  // the epilogue stage when not aliased
    unsigned char *const mring = est_own + S::EST;
    uint16_t *const thi = reinterpret_cast<uint16_t *>(mring + S::MRING);
    uint16_t *const tlo = thi + TW;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int col = lane & 15, kg = lane >> 4;

    int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t strip = tile % g.nstrip;
    tile /= g.nstrip;
    const int64_t seg = tile % g.nseg;
    const int64_t z = tile / g.nseg;
    const int64_t ny = g.ny, nx = g.nx;
    const int64_t x0 = strip * TX;
    const int64_t ya = seg * g.seg;
    const int64_t yb = ya + g.seg < ny ? ya + g.seg : ny;
    const int npairs = (int)((yb - ya + 31) / 32);
    const int nsteps = 2 * (npairs + W - 1);
    const uint16_t *const in_z = in + z * ny * nx;

    lowp_fill_taps<RT>(thi, tlo, taps);

    // ---- step loads: lane slot lq = (n NW + wv) 64 + lane of a plane's
    // staging is row lq / PR, LDS piece lq % PR, which holds the global piece
    // (lq % PR) ^ (row & kXYSwz) of that row (swizzled so that the x-pass
    // operand reads of consecutive rows fall in distinct bank groups).  The row and
    // column of each of the lane's loads are fixed: only the step's first row
    // moves.
    constexpr int NW = S::NW;
    const uint32_t stage_lds = lds_addr(staging);
    const uint32_t mring_lds = lds_addr(mring);
    int lrow[NI];
    int64_t lcol[NI];
    bool lcol_ok[NI];
#pragma unroll
    for (int n = 0; n < NI; ++n) {
        const int lq = (n * NW + wv) * 64 + lane;
        const int pc = (lq % PR) ^ ((lq / PR) & kXYSwz);
        lrow[n] = lq / PR;
        lcol[n] = x0 - S::HALO + 8 * pc;
        lcol_ok[n] = pc >= S::PLO && pc < PR - S::PLO && lcol[n] >= 0 && lcol[n] < nx;
    }
    const int mw = 4 * wv + (lane >> 4);  // mask word column of this lane (waves < MLOADS)
    const bool mcol_ok = wv < S::MLOADS && mw < MW && x0 + 32 * mw < nx;
    const uint32_t *const mbase = bits + (z * g.wpr + x0 / 32 + mw) * ny;
    auto issue_group = [&](int i) {
        const int buf = i % (D + 1);
        const int64_t yrow0 = ya - RT + 16 * (int64_t)i;
        const bool live = i < nsteps;  // steps past the end load zeros (fixed counts)
#pragma unroll
        for (int n = 0; n < NI; ++n) {
            const int64_t y = yrow0 + lrow[n];
            const bool ok = live && lcol_ok[n] && y >= 0 && y < ny;
            const char *src = ok ? reinterpret_cast<const char *>(in_z + y * nx + lcol[n])
                                 : reinterpret_cast<const char *>(g_dma_zero);
            const uint32_t dst = stage_lds + (uint32_t)((buf * P) * SB + (n * NW + wv) * 1024);
            dma16<false>(dst, src);
            if constexpr (SPLIT)
                dma16<false>(dst + SB, ok ? src + 2 * io.in_lo : src);
        }
        // one mask load per wave: half h = i % 2 of pair q = i / 2 - (W - 1),
        // word column 4 wv + lane / 16, row 16 h + lane % 16; the other waves
        // load zeros into a spare target
        const int q = i / 2 - (W - 1), h = i & 1;
        const int64_t yy = ya + 32 * (int64_t)q + 16 * h + (lane & 15);
        const bool mine = wv < S::MLOADS && q >= 0;
        const bool mok = live && mine && mcol_ok && yy < ny;
        const void *msrc = mok ? static_cast<const void *>(mbase + yy) : static_cast<const void *>(g_dma_zero);
        const uint32_t mdst = mine ? mring_lds + (uint32_t)((q % (D + 1)) * MSLOT + h * MHALF + wv * 256)
                                   : mring_lds + (uint32_t)((D + 1) * MSLOT);
        dma4(mdst, msrc);
    };

#pragma unroll
    for (int i = 0; i < D; ++i) issue_group(i);
    __syncthreads();  // tap tables
    v8s ahi[C], alo[C];
    lowp_read_taps<RT>(thi, tlo, col, kg, ahi, alo);

    // y-pass: wave wv takes quad wv % 4 of the sub-rings wv / 4 + 2 i; the
    // transposed-read offsets are the 16-bit body's: lane (kg, col) reads row
    // rl of a slot, lines 4 (col & 3) .. +3 of its quad
    const int quad = wv & 3;
    int roff[2][2], radj[2][2];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            const int rl = 16 * u + 8 * kg + (col >> 2) + 4 * hh;
            const int rr = rl & 31;
            radj[u][hh] = rl >> 5;
            roff[u][hh] = rr * 128 + ((quad ^ row_swizzle(rr)) << 5) + 8 * (col & 3);
        }
    uint16_t *const sink = reinterpret_cast<uint16_t *>(g_dma_sink + lane);
    float *const sinkf = reinterpret_cast<float *>(g_dma_sink + lane);

    for (int j = 0; j < nsteps; ++j) {
        // after step j's loads the wave issued D - 1 more steps' loads and the
        // stores of the epilogues among steps j - D .. j - 1 (odd steps from
        // 2 W - 1 on)
        int n_st = 0;
#pragma unroll
        for (int k = 1; k <= D; ++k) {
            const int jj = j - k;
            n_st += (jj >= 2 * W - 1 && (jj & 1)) ? 1 : 0;
        }
        if (n_st == 0)
            vm_wait<(D - 1) * GL>();
        else if (n_st == 1)
            vm_wait<(D - 1) * GL + ST>();
        else if (n_st == 2)
            vm_wait<(D - 1) * GL + (S::NST >= 2 ? 2 : 1) * ST>();
        else
            vm_wait<(D - 1) * GL + S::NST * ST>();
        __syncthreads();
        issue_group(j + D);
        unsigned char *const stg = staging + (j % (D + 1)) * P * SB;
        // the block epilogue of this (odd) step stages in stg when aliased:
        // every wave's x-pass reads of it precede the y-pass barrier
        unsigned char *const est = S::ALIAS && !S::WEPI ? stg : est_own;

        // ---- x-pass of the step's 16 rows.  Wave wv takes the column tiles
        // t0 + 2 i (i < NT), t0 = (wv & 1) + 2 NT (wv >> 1): tile t + 2
        // starts one 32-position chunk after tile t, so chunk c + i of one set
        // of C + NT - 1 operand reads feeds tile t0 + 2 i's chunk c.
        unsigned char *const slot = ring + ((j >> 1) % W) * SLOT;
        {
            constexpr int NT = S::NT;            // tiles per wave
            constexpr int NR = C + NT - 1;       // operand reads per wave
            const int r = col;                   // this lane's staged row (MFMA line)
            const int rs = 16 * (j & 1) + r;     // its row in the ring slot
            const unsigned char *const rowp = stg + r * PR * 16;
            const int t0 = (wv & 1) + 2 * NT * (wv >> 1);
            v4f acc[NT];
#pragma unroll
            for (int i = 0; i < NT; ++i) acc[i] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
            v8s xs[NR], xls[NR];
#pragma unroll
            for (int q = 0; q < NR; ++q) {
                // halves [o, o + 8) of the staged row
                const int o = S::HALO - RT + 16 * t0 + 32 * q + 8 * kg;
                v8s x, xl;
                if constexpr (RT % 8 == 0) {
                    // one whole 16-byte piece: conflict-free across the wave
                    const int boff = ((o >> 3) ^ (r & kXYSwz)) << 4;
                    x = *reinterpret_cast<const v8s *>(rowp + boff);
                    if constexpr (SPLIT) xl = *reinterpret_cast<const v8s *>(rowp + SB + boff);
                    if constexpr (S::GUARD) {
                        // past the row end (the read stayed inside the LDS:
                        // the next row, or the ring after the last buffer)
                        if (o >= S::ROWH) {
                            x = v8s{0, 0, 0, 0, 0, 0, 0, 0};
                            if constexpr (SPLIT) xl = x;
                        }
                    }
                } else {
                    static_assert(!S::GUARD, "overhang guard for whole 16-byte pieces only");
                    uint2 part[2], plo[2];
#pragma unroll
                    for (int hh = 0; hh < 2; ++hh) {
                        const int oh = o + 4 * hh;
                        const int boff = (((oh >> 3) ^ (r & kXYSwz)) << 4) + ((oh & 7) << 1);
                        part[hh] = *reinterpret_cast<const uint2 *>(rowp + boff);
                        if constexpr (SPLIT) plo[hh] = *reinterpret_cast<const uint2 *>(rowp + SB + boff);
                    }
                    x = __builtin_bit_cast(v8s, make_uint4(part[0].x, part[0].y, part[1].x, part[1].y));
                    if constexpr (SPLIT)
                        xl = __builtin_bit_cast(v8s, make_uint4(plo[0].x, plo[0].y, plo[1].x, plo[1].y));
                }
                xs[q] = x;
                if constexpr (SPLIT) xls[q] = xl;
                if constexpr (!S::PRELOAD) xy_x_products<L, NT, C, SPLIT>(q, xs, xls, ahi, alo, acc);
            }
            if constexpr (S::PRELOAD) {
                __builtin_amdgcn_sched_barrier(0);  // keeps the reads ahead of the products
#pragma unroll
                for (int q = 0; q < NR; ++q) xy_x_products<L, NT, C, SPLIT>(q, xs, xls, ahi, alo, acc);
            }
#pragma unroll
            for (int i = 0; i < NT; ++i) {
                // lane holds row r, columns 16 t + 4 kg .. +3: one 8-byte
                // piece of slot row rs in its sub-ring
                const int t = t0 + 2 * i;
                const int xc = 16 * t + 4 * kg;
                const int sub = xc >> 6, ln = xc & 63;
                const int boff = sub * kSlotBytes + rs * 128 + (((ln >> 4) ^ row_swizzle(rs)) << 5) + ((ln & 15) << 1);
                if constexpr (SPLIT) {
                    uint16_t hv[4], lv[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) split_f16_scaled(acc[i][e], io.scale, hv[e], lv[e]);
                    *reinterpret_cast<uint2 *>(slot + boff) = make_uint2(pack2(hv[0], hv[1]), pack2(hv[2], hv[3]));
                    *reinterpret_cast<uint2 *>(slot + W * SLOT + boff) =
                        make_uint2(pack2(lv[0], lv[1]), pack2(lv[2], lv[3]));
                } else {
                    *reinterpret_cast<uint2 *>(slot + boff) =
                        make_uint2(pack2(L::from_f32(acc[i][0]), L::from_f32(acc[i][1])),
                                   pack2(L::from_f32(acc[i][2]), L::from_f32(acc[i][3])));
                }
            }
        }
        if ((j & 1) == 0 || j < 2 * W - 1) continue;
        __syncthreads();

        // ---- y-pass of pair p from slots p .. p + W - 1
        const int p = (j >> 1) - (W - 1);
        const int sp = p % W;
        constexpr int NQ = S::NQ;
        v4f acc[NQ][2];
#pragma unroll
        for (int gq = 0; gq < NQ; ++gq)
#pragma unroll
            for (int u = 0; u < 2; ++u) acc[gq][u] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
        // operands of chunk c for group gq and row tile u: two transposed
        // 8-byte reads per plane (rows rl and rl + 4 of the slots)
        v8s yx[C][NQ][2], yxl[C][NQ][2];
        auto y_read = [&](int c, int gq, int u) {
            const int sb = (wv >> 2) + 2 * gq;  // sub-ring
            typedef __attribute__((address_space(3))) v4s lds_v4s;
            v4s part[2], plo[2];
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                int si = sp + c + radj[u][hh];
                si = si >= W ? si - W : si;
                si = si >= W ? si - W : si;
                const unsigned char *a = ring + si * SLOT + sb * kSlotBytes + roff[u][hh];
                part[hh] = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s *)a);
                if constexpr (SPLIT) plo[hh] = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s *)(a + W * SLOT));
            }
            yx[c][gq][u] = v8s{part[0][0], part[0][1], part[0][2], part[0][3],
                               part[1][0], part[1][1], part[1][2], part[1][3]};
            if constexpr (SPLIT)
                yxl[c][gq][u] = v8s{plo[0][0], plo[0][1], plo[0][2], plo[0][3],
                                    plo[1][0], plo[1][1], plo[1][2], plo[1][3]};
        };
        auto y_products = [&](int c, int gq, int u) {
            acc[gq][u] = L::mfma(ahi[c], yx[c][gq][u], acc[gq][u]);
            acc[gq][u] = L::mfma(alo[c], yx[c][gq][u], acc[gq][u]);
            if constexpr (SPLIT) acc[gq][u] = L::mfma(ahi[c], yxl[c][gq][u], acc[gq][u]);
        };
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int gq = 0; gq < NQ; ++gq)
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    y_read(c, gq, u);
                    if constexpr (!S::PRELOAD) y_products(c, gq, u);
                }
        if constexpr (S::PRELOAD) {
            __builtin_amdgcn_sched_barrier(0);  // keeps the reads ahead of the products
#pragma unroll
            for (int c = 0; c < C; ++c)
#pragma unroll
                for (int gq = 0; gq < NQ; ++gq)
#pragma unroll
                    for (int u = 0; u < 2; ++u) y_products(c, gq, u);
        }

        const unsigned char *const msl = mring + (p % (D + 1)) * MSLOT;
        if constexpr (S::WEPI) {
            // ---- epilogue, per wave: lane (kg, col) holds column
            // 64 sb + 16 quad + col of the strip, rows 16 u + 4 kg .. +3 of the
            // pair; the wave stages its 32 x 16 tile as [row][column] in its own
            // area, then lane l stores row l / 2, columns 8 (l % 2) .. +7 (the
            // same wave's next pair writes the area only after these reads)
            unsigned char *const west = est + wv * S::WEST;
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int row = 16 * u + 4 * kg + e;
                    if constexpr (FIN != 0) {
                        *reinterpret_cast<float *>(west + row * S::WPF + 4 * col) = acc[0][u][e] * io.scale;
                    } else if constexpr (SPLIT) {
                        uint16_t hv, lv;
                        split_f16_scaled(acc[0][u][e], io.scale, hv, lv);
                        *reinterpret_cast<uint16_t *>(west + row * S::WP16 + 2 * col) = hv;
                        *reinterpret_cast<uint16_t *>(west + 32 * S::WP16 + row * S::WP16 + 2 * col) = lv;
                    } else {
                        *reinterpret_cast<uint16_t *>(west + row * S::WP16 + 2 * col) = L::from_f32(acc[0][u][e]);
                    }
                }
            wave_lds_sync();
            const int rw = lane >> 1, hf = lane & 1;
            const int pc = 8 * (wv >> 2) + 2 * quad + hf;  // 8-column piece of the strip
            const int64_t y = ya + 32 * (int64_t)p + rw;
            const int64_t x = x0 + 8 * pc;
            const bool ok = y < yb && x < nx;
            const int64_t off = (z * ny + y) * nx + x;
            const uint32_t m =
                *reinterpret_cast<const uint32_t *>(msl + (rw >> 4) * MHALF + (pc >> 2) * 64 + (rw & 15) * 4) >>
                (8 * (pc & 3));
            if constexpr (FIN != 0) {
                v4f v[2];
#pragma unroll
                for (int hh = 0; hh < 2; ++hh)
                    v[hh] = *reinterpret_cast<const v4f *>(west + rw * S::WPF + 32 * hf + 16 * hh);
                xy_store_fin<FIN>(v, m, ok, off, io, sinkf, sink);
            } else {
                const uint4 v = *reinterpret_cast<const uint4 *>(west + rw * S::WP16 + 16 * hf);
                uint4 vl = make_uint4(0, 0, 0, 0);
                if constexpr (SPLIT) vl = *reinterpret_cast<const uint4 *>(west + 32 * S::WP16 + rw * S::WP16 + 16 * hf);
                xy_store_state<BF, SPLIT>(v, vl, m, ok, off, out, io, sink);
            }
            continue;
        }
        // ---- epilogue: lane (kg, col) of group gq holds column
        // 64 sb + 16 quad + col, rows 16 u + 4 kg .. +3 of the pair; stage as
        // [row][column] (the previous pair's stores read the stage before
        // this step's first barrier)
#pragma unroll
        for (int gq = 0; gq < NQ; ++gq)
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int row = 16 * u + 4 * kg + e;
                    const int xc = 64 * ((wv >> 2) + 2 * gq) + 16 * quad + col;
                    if constexpr (FIN != 0) {
                        *reinterpret_cast<float *>(est + stage_off<S::ALIAS>(row, 4 * xc, S::SPITCHF)) =
                            acc[gq][u][e] * io.scale;
                    } else if constexpr (SPLIT) {
                        uint16_t hv, lv;
                        split_f16_scaled(acc[gq][u][e], io.scale, hv, lv);
                        const int so = stage_off<S::ALIAS>(row, 2 * xc, S::SPITCH16);
                        *reinterpret_cast<uint16_t *>(est + so) = hv;
                        *reinterpret_cast<uint16_t *>(est + 32 * S::SPITCH16 + so) = lv;
                    } else {
                        *reinterpret_cast<uint16_t *>(est + stage_off<S::ALIAS>(row, 2 * xc, S::SPITCH16)) =
                            L::from_f32(acc[gq][u][e]);
                    }
                }
        __syncthreads();
        // 16-bit pieces: piece pc of row rw covers columns 8 pc .. +7
#pragma unroll
        for (int k = 0; k < S::PIECES16; ++k) {
            const int idx = k * 64 * NW + tid;
            const int rw = idx / (TX / 8), pc = idx % (TX / 8);
            const int64_t y = ya + 32 * (int64_t)p + rw;
            const int64_t x = x0 + 8 * pc;
            const bool ok = y < yb && x < nx;
            const int64_t off = (z * ny + y) * nx + x;
            const uint32_t m =
                *reinterpret_cast<const uint32_t *>(msl + (rw >> 4) * MHALF + (pc >> 2) * 64 + (rw & 15) * 4) >>
                (8 * (pc & 3));
            if constexpr (FIN != 0) {
                v4f v[2];
#pragma unroll
                for (int hh = 0; hh < 2; ++hh)
                    v[hh] = *reinterpret_cast<const v4f *>(est + stage_off<S::ALIAS>(rw, 32 * pc + 16 * hh, S::SPITCHF));
                xy_store_fin<FIN>(v, m, ok, off, io, sinkf, sink);
            } else {
                const int so = stage_off<S::ALIAS>(rw, 16 * pc, S::SPITCH16);
                const uint4 v = *reinterpret_cast<const uint4 *>(est + so);
                uint4 vl = make_uint4(0, 0, 0, 0);
                if constexpr (SPLIT) vl = *reinterpret_cast<const uint4 *>(est + 32 * S::SPITCH16 + so);
                xy_store_state<BF, SPLIT>(v, vl, m, ok, off, out, io, sink);
            }
        }
    }
    vm_wait<0>();
}

// ---------------------------------------------------------------- host side
// Strip width and prefetch depth per form (measured on 1024^3, R=40).
#ifndef MXB_XY_TX16
#define MXB_XY_TX16 128
#endif
#ifndef MXB_XY_D16
#define MXB_XY_D16 1
#endif
#ifndef MXB_XY_TXS
#define MXB_XY_TXS 128
#endif
#ifndef MXB_XY_DS
#define MXB_XY_DS 3
#endif

// Strip width and prefetch depth of a form and radius bucket: the deepest
// prefetch up to the form's default whose shape fits the LDS.  FITS: the
// fused pass serves the bucket.  The 16-bit forms take it where two blocks
// share a CU (80 KB of LDS each): with one block per CU it was slower than
// the three-pass form at buckets 48 and 64 (1024^3 BF16, sigma 12: 251.7
// against 260.0 Gvoxel*iter/s); with the aliased stage and one step of
// prefetch every bucket fits two or more, and sigma 12 / 15 run 281.0 / 256.0
// against 260.6 / 240.7.  The split form gains at every bucket (sigma 12:
// 195.2 against 159.8; sigma 15: 160.6 against 135.4).
template <int RT, bool SPLIT>
struct XYForm {
    static constexpr int TX = SPLIT ? MXB_XY_TXS : MXB_XY_TX16;
    static constexpr int D0 = SPLIT ? MXB_XY_DS : MXB_XY_D16;
    static constexpr int D = XYShape<RT, TX, D0, SPLIT, 0>::FITS ? D0
                             : (D0 > 2 && XYShape<RT, TX, 2, SPLIT, 0>::FITS ? 2 : 1);
    static constexpr bool FITS = XYShape<RT, TX, D, SPLIT, 0>::FITS &&
                                 XYShape<RT, TX, D, SPLIT, SPLIT ? 3 : 0>::FITS &&
                                 (SPLIT || XYShape<RT, TX, D, SPLIT, 0>::BYTES <= 80 * 1024);
};

// Whether the form's kernel for radius R fits the LDS.
template <bool SPLIT>
static bool xy_fits(int R) {
#define MXB_XY_FITS(rt)                                                                                  \
    if (R <= rt) return XYForm<rt, SPLIT>::FITS;
    MXB_XY_FITS(4)
    MXB_XY_FITS(8)
    MXB_XY_FITS(12)
    MXB_XY_FITS(16)
    MXB_XY_FITS(24)
    MXB_XY_FITS(32)
    MXB_XY_FITS(40)
    MXB_XY_FITS(48)
    MXB_XY_FITS(64)
#undef MXB_XY_FITS
    return false;
}

// Usable when rows are whole 16-byte pieces and 16-byte aligned, the radius
// is compiled, the kernel fits the LDS and the mask bits are there.
bool xy_fused_supported(const uint16_t *in, const uint16_t *out, const MaskBits &bits, int64_t nz, int64_t ny,
                        int64_t nx, int R, int precision, const SplitIO &io) {
    if (R > kMaxRadiusLowp || nz * ny * nx == 0 || nx % 8 != 0 || bits.w == nullptr) return false;
    if (!(precision == MXB_FP32_SPLIT ? xy_fits<true>(R) : xy_fits<false>(R))) return false;
    if (precision != MXB_FP32_SPLIT && (io.fin_f32 != nullptr || io.fin_u16 != nullptr)) return false;
    if (reinterpret_cast<uintptr_t>(in) % 16 != 0 || reinterpret_cast<uintptr_t>(out) % 16 != 0) return false;
    if (io.in_lo % 8 != 0 || io.out_lo % 8 != 0) return false;
    if (io.fin_f32 && reinterpret_cast<uintptr_t>(io.fin_f32) % 16 != 0) return false;
    if (io.fin_u16 && reinterpret_cast<uintptr_t>(io.fin_u16) % 16 != 0) return false;
    return true;
}

template <int RT, bool BF, bool SPLIT>
static hipError_t launch_xy_rt(const uint16_t *in, uint16_t *out, const MaskBits &bits, int64_t nz, int64_t ny,
                               int64_t nx, const float *k, int R, SplitIO io, hipStream_t s) {
    constexpr int TX = XYForm<RT, SPLIT>::TX;
    constexpr int D = XYForm<RT, SPLIT>::D;
    XYGeom g;
    g.nz = nz;
    g.ny = ny;
    g.nx = nx;
    g.nstrip = (nx + TX - 1) / TX;
    g.wpr = mask_words_per_row(nx);
    // whole columns per block unless the grid would be too small to fill the chip
    const int64_t cols = nz * g.nstrip;
    int64_t nseg = (256 * 2 + cols - 1) / cols;
    const int64_t max_seg = (ny + 63) / 64;
    if (nseg > max_seg) nseg = max_seg;
    if (nseg < 1) nseg = 1;
    g.seg = ((ny + nseg - 1) / nseg + 31) / 32 * 32;
    g.nseg = (ny + g.seg - 1) / g.seg;
    const int64_t blocks = g.nseg * cols;
    if (blocks == 0) return hipSuccess;
    TapsL<2 * RT + 1> t;
    if (SPLIT) {
        float scale = 1.0f;
        t = split_taps_scaled<RT>(k, R, &scale);
        io.scale = scale;
    } else {
        t = split_taps<RT>(k, R, BF);
    }
    const int fin = (io.fin_f32 != nullptr ? 1 : 0) + (io.fin_u16 != nullptr ? 2 : 0);
    const dim3 grid((unsigned)blocks), block(64 * kXYWaves);
    const uint32_t *w = bits.w;
    if constexpr (!XYForm<RT, SPLIT>::FITS) {
        return hipErrorInvalidValue;
    } else if constexpr (SPLIT) {
        switch (fin) {
        case 0: hipLaunchKernelGGL((k_xy_fused<RT, TX, D, false, true, 0>), grid, block, 0, s, in, out, w, g, t, io); break;
        case 1: hipLaunchKernelGGL((k_xy_fused<RT, TX, D, false, true, 1>), grid, block, 0, s, in, out, w, g, t, io); break;
        case 2: hipLaunchKernelGGL((k_xy_fused<RT, TX, D, false, true, 2>), grid, block, 0, s, in, out, w, g, t, io); break;
        default: hipLaunchKernelGGL((k_xy_fused<RT, TX, D, false, true, 3>), grid, block, 0, s, in, out, w, g, t, io); break;
        }
    } else {
        if (fin != 0) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_xy_fused<RT, TX, D, BF, false, 0>), grid, block, 0, s, in, out, w, g, t, io);
    }
    return hipGetLastError();
}

// precision: MXB_BF16 / MXB_FP16 (one 16-bit plane) or MXB_FP32_SPLIT (two
// FP16 planes, io.*_lo; final outputs when io.fin_* are set).
hipError_t launch_xy_fused(const uint16_t *in, uint16_t *out, const MaskBits &bits, int64_t nz, int64_t ny,
                           int64_t nx, const float *taps, int R, int precision, const SplitIO &io, hipStream_t s) {
    if (!xy_fused_supported(in, out, bits, nz, ny, nx, R, precision, io)) return hipErrorInvalidValue;
    const bool split = precision == MXB_FP32_SPLIT;
#define MXB_XY_CASE(rt)                                                                             \
    if (R <= rt) {                                                                                  \
        if (split) return launch_xy_rt<rt, false, true>(in, out, bits, nz, ny, nx, taps, R, io, s); \
        if (precision == MXB_BF16) return launch_xy_rt<rt, true, false>(in, out, bits, nz, ny, nx, taps, R, io, s); \
        return launch_xy_rt<rt, false, false>(in, out, bits, nz, ny, nx, taps, R, io, s);            \
    }
    // the 16-bit forms' radius buckets
    MXB_XY_CASE(4)
    MXB_XY_CASE(8)
    MXB_XY_CASE(12)
    MXB_XY_CASE(16)
    MXB_XY_CASE(24)
    MXB_XY_CASE(32)
    MXB_XY_CASE(40)
    MXB_XY_CASE(48)
    MXB_XY_CASE(64)
#undef MXB_XY_CASE
    return hipErrorInvalidValue;
}

}  // namespace mxb
