// Consumers of the diffusion field (gfx950): field combination and 2D
// histograms.  C ABI in include/maxibone_fields.h.
//
// Layout: all volumes are C-order (z, y, x); a "row" is one (z, y) line of
// nx voxels.  Every kernel walks rows with 128-thread blocks, 8 consecutive
// voxels per thread (one 16-byte load of uint16 where the row is aligned),
// so a block covers 1024 voxels of a row.  All the work is HBM-bound byte
// arithmetic (2-4 bytes read per voxel); nothing here is GEMM-shaped.
//
// Combination (0975_combine_fields.py:41-62, float32 in the script's order):
//   c = g/gmax - e/emax;  c' = c + |min c|;  w = c' / max c';  v = w * cyl;
//   out = cyl * (uint16)((v - min v) / (max v - min v) * 65535)
// Rounding is monotone, so max c' = fl(max c + a), and min/max of v follow
// from min/max of c over the cylinder; three passes therefore reproduce the
// script exactly: (1) max g, max e; (2) min/max of c overall and inside the
// cylinder; (3) the output.
//
// Histograms (cpu/histograms.cc): the float64 bin expressions depend only on
// the uint16 value, so the host evaluates them once per value into lookup
// tables (-1 = not counted); field coordinates floor(Z*nz/nZ) etc. become
// per-axis maps.  Kernels add with 64-bit global atomics (vector memory
// atomics), after which the caller's bins hold the sums.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "maxibone_diffusion.h"
#include "maxibone_fields.h"
#include "mxb_internal.h"

namespace mxb {

constexpr int kRowThreads = 128;
constexpr int kPerThread = 8;
constexpr int kRowChunk = kRowThreads * kPerThread;  // voxels of a row per block

struct RowWalk {
    int64_t rows, nx, chunks;  // chunks per row
};

static RowWalk row_walk(int64_t rows, int64_t nx) {
    return RowWalk{rows, nx, (nx + kRowChunk - 1) / kRowChunk};
}

// Load 8 consecutive uint16 of a row starting at x0 (elements past nx read 0).
__device__ __forceinline__ void load8(const uint16_t *__restrict__ row, int64_t x0, int64_t nx,
                                      bool vec, uint16_t v[8]) {
    if (vec && x0 + 8 <= nx) {
        const uint4 q = *reinterpret_cast<const uint4 *>(row + x0);
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = (uint16_t)(w[k >> 1] >> (16 * (k & 1)));
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = x0 + k < nx ? row[x0 + k] : (uint16_t)0;
    }
}

// Orderable encoding of a float for atomic min / max on uint32.
__device__ __forceinline__ uint32_t f2ord(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
static float ord2f(uint32_t o) {
    const uint32_t b = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}
__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}

// ws[0] = max g, ws[1] = max e, ws[2..5] = ord(min c), ord(max c),
// ord(min c | cyl), ord(max c | cyl).
__global__ __launch_bounds__(64) void k_fields_init(uint32_t *ws) {
    if (threadIdx.x == 0) {
        ws[0] = 0; ws[1] = 0;
        ws[2] = 0xffffffffu; ws[3] = 0u;
        ws[4] = 0xffffffffu; ws[5] = 0u;
    }
}

__global__ __launch_bounds__(kRowThreads) void k_max2(const uint16_t *__restrict__ a,
                                                      const uint16_t *__restrict__ b, RowWalk w,
                                                      int vec, uint32_t *ws) {
    uint32_t ma = 0, mb = 0;
    // grid-stride over (row, chunk): one atomic per block at the end
    for (int64_t i = blockIdx.x; i < w.rows * w.chunks; i += gridDim.x) {
        const int64_t row = i / w.chunks;
        const int64_t x0 = (i % w.chunks) * kRowChunk + threadIdx.x * kPerThread;
        if (x0 >= w.nx) continue;
        uint16_t va[8], vb[8];
        load8(a + row * w.nx, x0, w.nx, vec != 0, va);
        load8(b + row * w.nx, x0, w.nx, vec != 0, vb);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            ma = max(ma, (uint32_t)va[k]);
            mb = max(mb, (uint32_t)vb[k]);
        }
    }
    __shared__ uint32_t red[2][kRowThreads / 64];
    ma = wave_max(ma);
    mb = wave_max(mb);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = ma;
        red[1][threadIdx.x >> 6] = mb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kRowThreads / 64; ++i) {
            ma = max(ma, red[0][i]);
            mb = max(mb, red[1][i]);
        }
        atomicMax(&ws[0], ma);
        atomicMax(&ws[1], mb);
    }
}

__device__ __forceinline__ float combine_c(uint16_t g, uint16_t e, float gm, float em) {
#pragma clang fp contract(off)
    return (float)g / gm - (float)e / em;
}

__global__ __launch_bounds__(kRowThreads) void k_combine_minmax(const uint16_t *__restrict__ g,
                                                                const uint16_t *__restrict__ e,
                                                                const uint8_t *__restrict__ cyl,
                                                                RowWalk w, int64_t ny, int vec,
                                                                float gm, float em, uint32_t *ws) {
    uint32_t lo = 0xffffffffu, hi = 0u, lo_c = 0xffffffffu, hi_c = 0u;
    for (int64_t i = blockIdx.x; i < w.rows * w.chunks; i += gridDim.x) {
        const int64_t row = i / w.chunks;
        const int64_t x0 = (i % w.chunks) * kRowChunk + threadIdx.x * kPerThread;
        if (x0 >= w.nx) continue;
        uint16_t vg[8], ve[8];
        load8(g + row * w.nx, x0, w.nx, vec != 0, vg);
        load8(e + row * w.nx, x0, w.nx, vec != 0, ve);
        const uint8_t *crow = cyl != nullptr ? cyl + (row % ny) * w.nx : nullptr;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (x0 + k >= w.nx) break;
            const uint32_t o = f2ord(combine_c(vg[k], ve[k], gm, em));
            lo = min(lo, o);
            hi = max(hi, o);
            if (crow == nullptr || crow[x0 + k] != 0) {
                lo_c = min(lo_c, o);
                hi_c = max(hi_c, o);
            }
        }
    }
    __shared__ uint32_t red[4][kRowThreads / 64];
    lo = wave_min(lo); hi = wave_max(hi); lo_c = wave_min(lo_c); hi_c = wave_max(hi_c);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = lo; red[1][threadIdx.x >> 6] = hi;
        red[2][threadIdx.x >> 6] = lo_c; red[3][threadIdx.x >> 6] = hi_c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kRowThreads / 64; ++i) {
            lo = min(lo, red[0][i]); hi = max(hi, red[1][i]);
            lo_c = min(lo_c, red[2][i]); hi_c = max(hi_c, red[3][i]);
        }
        atomicMin(&ws[2], lo); atomicMax(&ws[3], hi);
        atomicMin(&ws[4], lo_c); atomicMax(&ws[5], hi_c);
    }
}

struct CombineParams {
    float gm, em, a, M, vmin, range;
};

__global__ __launch_bounds__(kRowThreads) void k_combine_out(const uint16_t *__restrict__ g,
                                                             const uint16_t *__restrict__ e,
                                                             const uint8_t *__restrict__ cyl,
                                                             RowWalk w, int64_t ny, int vec,
                                                             CombineParams p,
                                                             uint16_t *__restrict__ out) {
#pragma clang fp contract(off)
    const int64_t row = blockIdx.x / w.chunks;
    const int64_t x0 = (blockIdx.x % w.chunks) * kRowChunk + threadIdx.x * kPerThread;
    if (row >= w.rows || x0 >= w.nx) return;
    uint16_t vg[8], ve[8], q[8];
    load8(g + row * w.nx, x0, w.nx, vec != 0, vg);
    load8(e + row * w.nx, x0, w.nx, vec != 0, ve);
    const uint8_t *crow = cyl != nullptr ? cyl + (row % ny) * w.nx : nullptr;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const bool inside = crow == nullptr || x0 + k >= w.nx || crow[x0 + k] != 0;
        const float c = combine_c(vg[k], ve[k], p.gm, p.em);
        const float wv = (c + p.a) / p.M;
        const float v = inside ? wv : 0.0f;                 // combined *= cylinder_mask
        const float s = (v - p.vmin) / p.range * 65535.0f;  // general.hh:100
        const uint16_t u = p.range > 0.0f ? (uint16_t)(uint32_t)s : (uint16_t)0;
        q[k] = inside ? u : (uint16_t)0;                     // combined *= cylinder_mask
    }
    uint16_t *orow = out + row * w.nx;
    if (vec != 0 && x0 + 8 <= w.nx) {
        *reinterpret_cast<uint4 *>(orow + x0) =
            make_uint4((uint32_t)q[0] | ((uint32_t)q[1] << 16), (uint32_t)q[2] | ((uint32_t)q[3] << 16),
                       (uint32_t)q[4] | ((uint32_t)q[5] << 16), (uint32_t)q[6] | ((uint32_t)q[7] << 16));
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (x0 + k < w.nx) orow[x0 + k] = q[k];
    }
}

// ------------------------------------------------------------- histograms
struct FieldHistArgs {
    int64_t nZ, nY, nX, nz, ny, nx, vb, fb;
    const int32_t *vlut, *flut;          // per uint16 value: bin or -1
    const int32_t *zmap, *ymap, *xmap;   // voxel coordinate -> field coordinate
};

// Field histogram.  The volume is walked in bricks of kBrickZ x kBrickY x
// kBrickX voxels; each block takes one contiguous range of bricks.  A brick
// covers a small part of the (smooth) field, so few field bins, and its
// (field bin, voxel bin) pairs repeat: they gather in an LDS hash table of
// kHashSlots keys (linear probing) that also counts the distinct keys it
// holds.  The table reaches the global bins, one atomic per distinct pair,
// when it holds more than kFlushAt keys after a step of kRowChunk voxels, and
// at the end of the range; a step adds at most kRowChunk keys, so the table
// is never more than half full and a free slot or the key itself is always
// found.  Global atomics are the cost here: walking whole rows, which cross
// the whole field, took 127 ms on 1024^3 (4096 x 2048 bins), of which 20.7 ms
// without the atomics.
constexpr int kHashSlots = 8192;
constexpr int kHashShift = 19;  // 32 - log2(kHashSlots)
constexpr uint32_t kFlushAt = kHashSlots / 2 - kRowChunk;
constexpr unsigned kFieldHistBlocks = 512;  // two 64 KB blocks per CU of a 256-CU device
constexpr int kBrickX = 64, kBrickY = 32, kBrickZ = 32;
constexpr int kBrickRowsPerStep = kRowChunk / kBrickX;  // 16 brick rows of 64 voxels per step
constexpr uint32_t kEmpty = 0xffffffffu;
static_assert(kHashSlots == (1 << (32 - kHashShift)), "hash table size");
static_assert(kBrickX % kPerThread == 0 && (kBrickY * kBrickZ) % kBrickRowsPerStep == 0, "brick shape");

__global__ __launch_bounds__(kRowThreads) void k_field_histogram(const uint16_t *__restrict__ vox, const uint16_t *__restrict__ field,
                                                                 FieldHistArgs a, int vec,
                                                                 unsigned long long *__restrict__ bins) {
    __shared__ uint32_t keys[kHashSlots];
    __shared__ uint32_t counts[kHashSlots];
    __shared__ uint32_t n_keys;
    for (int i = threadIdx.x; i < kHashSlots; i += kRowThreads) {
        keys[i] = kEmpty;
        counts[i] = 0;
    }
    if (threadIdx.x == 0) n_keys = 0;
    __syncthreads();
    const int64_t bx_n = (a.nX + kBrickX - 1) / kBrickX, by_n = (a.nY + kBrickY - 1) / kBrickY;
    const int64_t bricks = bx_n * by_n * ((a.nZ + kBrickZ - 1) / kBrickZ);
    const int64_t per = (bricks + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = (int64_t)blockIdx.x * per, b1 = b0 + per < bricks ? b0 + per : bricks;
    const int xo = (threadIdx.x % (kBrickX / kPerThread)) * kPerThread;
    const int ro = threadIdx.x / (kBrickX / kPerThread);
    constexpr int kSteps = kBrickY * kBrickZ / kBrickRowsPerStep;
    for (int64_t b = b0; b < b1; ++b) {
        const int64_t bx = b % bx_n, by = (b / bx_n) % by_n, bz = b / (bx_n * by_n);
        const int64_t x0 = bx * kBrickX + xo;
        for (int step = 0; step < kSteps; ++step) {
            const int r = step * kBrickRowsPerStep + ro;
            const int64_t Y = by * kBrickY + r % kBrickY, Z = bz * kBrickZ + r / kBrickY;
            if (x0 < a.nX && Y < a.nY && Z < a.nZ) {
                uint16_t v[8];
                load8(vox + (Z * a.nY + Y) * a.nX, x0, a.nX, vec != 0, v);
                const uint16_t *frow = field + ((int64_t)a.zmap[Z] * a.ny + a.ymap[Y]) * a.nx;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    if (x0 + k >= a.nX) break;
                    const int32_t vi = a.vlut[v[k]];
                    if (vi < 0) continue;
                    const int32_t fi = a.flut[frow[a.xmap[x0 + k]]];
                    if (fi < 0) continue;
                    const uint32_t key = (uint32_t)fi * (uint32_t)a.vb + (uint32_t)vi;
                    uint32_t slot = (key * 2654435761u) >> kHashShift;
                    while (true) {
                        const uint32_t prev = atomicCAS(&keys[slot], kEmpty, key);
                        if (prev == kEmpty || prev == key) {
                            atomicAdd(&counts[slot], 1u);
                            if (prev == kEmpty) atomicAdd(&n_keys, 1u);
                            break;
                        }
                        slot = (slot + 1) & (kHashSlots - 1);
                    }
                }
            }
            __syncthreads();
            const bool flush = n_keys > kFlushAt || (b + 1 == b1 && step + 1 == kSteps);  // uniform: between two barriers
            __syncthreads();
            if (flush) {
                for (int i = threadIdx.x; i < kHashSlots; i += kRowThreads) {
                    const uint32_t key = keys[i];
                    if (key != kEmpty) {
                        atomicAdd(&bins[key], (unsigned long long)counts[i]);
                        keys[i] = kEmpty;
                        counts[i] = 0;
                    }
                }
                if (threadIdx.x == 0) n_keys = 0;
                __syncthreads();
            }
        }
    }
}

struct AxisHistArgs {
    int64_t nz, ny, nx, z_offset, Nz, Nr, vb;
    const int32_t *vlut;  // per uint16 value: bin or -1
    int64_t cy, cx;
};

__device__ __forceinline__ int64_t isqrt64(int64_t s) {
    int64_t r = (int64_t)sqrt((double)s);
    while (r * r > s) --r;
    while ((r + 1) * (r + 1) <= s) ++r;
    return r;
}

// Axis histograms, y and z bins.  A block walks whole rows (grid-stride); the
// voxels of a row share y and z, so their bin counts gather in LDS (one
// histogram for the row's y, one for the current z plane) and reach y_bins /
// z_bins as one atomic per nonzero bin: at the end of each row for y, when
// the plane changes (rows of a block come in increasing order) for z.  x and
// r vary along the row; k_axis_xr counts them.  LDS: 2 x vb uint32.
__global__ __launch_bounds__(kRowThreads) void k_axis_histograms(const uint16_t *__restrict__ vox,
                                                                 RowWalk w, AxisHistArgs a, int vec,
                                                                 unsigned long long *__restrict__ yb,
                                                                 unsigned long long *__restrict__ zb) {
    extern __shared__ uint32_t hist[];  // [0, vb): row (y); [vb, 2 vb): plane (z)
    uint32_t *hy = hist, *hz = hist + a.vb;
    for (int64_t i = threadIdx.x; i < 2 * a.vb; i += kRowThreads) hist[i] = 0;
    __syncthreads();
    int64_t zcur = -1;  // plane whose counts hz holds (uniform)
    auto flush = [&](uint32_t *h, unsigned long long *dst) {
        for (int64_t i = threadIdx.x; i < a.vb; i += kRowThreads) {
            const uint32_t c = h[i];
            if (c != 0) {
                atomicAdd(&dst[i], (unsigned long long)c);
                h[i] = 0;
            }
        }
    };
    for (int64_t row = blockIdx.x; row < w.rows; row += gridDim.x) {
        const int64_t y = row % a.ny, z = row / a.ny + a.z_offset;
        if (z >= a.Nz) break;  // rows come in increasing order: the rest lie beyond Nz too
        if (z != zcur) {
            if (zcur >= 0) {
                __syncthreads();
                flush(hz, zb + zcur * a.vb);
                __syncthreads();
            }
            zcur = z;
        }
        for (int64_t x0 = (int64_t)threadIdx.x * kPerThread; x0 < w.nx; x0 += kRowChunk) {
            uint16_t v[8];
            load8(vox + row * w.nx, x0, w.nx, vec != 0, v);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int64_t x = x0 + k;
                if (x >= w.nx) break;
                const int32_t vi = a.vlut[v[k]];
                if (vi < 0) continue;
                atomicAdd(&hy[vi], 1u);
                atomicAdd(&hz[vi], 1u);
            }
        }
        __syncthreads();
        flush(hy, yb + y * a.vb);
        __syncthreads();
    }
    if (zcur >= 0) {
        __syncthreads();
        flush(hz, zb + zcur * a.vb);
    }
}

// Lanes of a wave that hold the same key (among those with `valid`) are
// combined: returns true on one lane per distinct key, with `count` the
// number of lanes that hold it.  After kMaxGroups keys the remaining lanes
// return true with count 1 (bounded work when the keys are mostly distinct:
// on noisy CT data 0-4 groups cost the same, 16 cost 8 % more, 64 2.2x).
// Call from wave-uniform control flow.
constexpr int kMaxGroups = 4;
__device__ __forceinline__ bool wave_group(uint32_t key, bool valid, uint32_t &count) {
    const int lane = (int)__lane_id();
    uint64_t pending = __ballot(valid);
    bool leader = false;
    count = 0;
    for (int it = 0; it < kMaxGroups && pending != 0; ++it) {
        const int l = __ffsll((unsigned long long)pending) - 1;
        const uint32_t k = (uint32_t)__shfl((int)key, l);
        const uint64_t same = __ballot(valid && key == k) & pending;
        if (lane == l) {
            leader = true;
            count = (uint32_t)__popcll(same);
        }
        pending &= ~same;
    }
    if ((pending >> lane) & 1ull) {
        leader = true;
        count = 1;
    }
    return leader;
}

// Axis histograms, x and r bins.  Counted along rows, the lanes would hold
// different x, so each voxel cost one global atomic for x and one for r
// (45-52 ms each on 1024^3 with 4096 bins).  Here a block takes COLS columns
// x0 .. x0+COLS-1 and kXrThreads consecutive z (one per lane) and walks y:
// the lanes of a wave then share x, y and therefore r, and differ only in the
// voxel bin.  x counts gather in LDS (COLS x vb uint32, at most kXrLds
// bytes: 8 columns up to 4096 bins, 4 up to 8192) and reach x_bins once per
// block; r counts are combined over the wave (one atomic per distinct voxel
// bin) and added to r_bins.
constexpr int kXrThreads = 256;
constexpr int64_t kXrLds = 128 * 1024;

// COLS (4 or 8) consecutive uint16 of a row starting at x0 (past nx read 0).
template <int COLS>
__device__ __forceinline__ void load_cols(const uint16_t *__restrict__ row, int64_t x0, int64_t nx, bool vec,
                                          uint16_t v[8]) {
    if constexpr (COLS == 8) {
        load8(row, x0, nx, vec, v);
    } else if (vec && x0 + 4 <= nx) {
        const uint2 q = *reinterpret_cast<const uint2 *>(row + x0);
        const uint32_t w[2] = {q.x, q.y};
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = (uint16_t)(w[k >> 1] >> (16 * (k & 1)));
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = x0 + k < nx ? row[x0 + k] : (uint16_t)0;
    }
}

template <int COLS>
__global__ __launch_bounds__(kXrThreads) void k_axis_xr(const uint16_t *__restrict__ vox, AxisHistArgs a,
                                                        int vec, unsigned long long *__restrict__ xb,
                                                        unsigned long long *__restrict__ rb) {
    constexpr int kXrCols = COLS;
    extern __shared__ uint32_t hx[];  // [COLS][vb]
    for (int64_t i = threadIdx.x; i < kXrCols * a.vb; i += kXrThreads) hx[i] = 0;
    __syncthreads();
    const int64_t ncol = (a.nx + kXrCols - 1) / kXrCols;
    const int64_t x0 = (blockIdx.x % ncol) * kXrCols;
    const int64_t z = (blockIdx.x / ncol) * kXrThreads + threadIdx.x;
    const bool zok = z < a.nz && z + a.z_offset < a.Nz;
    const uint16_t *base = vox + (zok ? z : 0) * a.ny * a.nx;
    uint16_t cur[8] = {0, 0, 0, 0, 0, 0, 0, 0}, nxt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (zok && a.ny > 0) load_cols<COLS>(base, x0, a.nx, vec != 0, cur);
    for (int64_t y = 0; y < a.ny; ++y) {
        if (zok && y + 1 < a.ny) load_cols<COLS>(base + (y + 1) * a.nx, x0, a.nx, vec != 0, nxt);
        const int64_t dy2 = (y - a.cy) * (y - a.cy);
#pragma unroll
        for (int k = 0; k < kXrCols; ++k) {
            const int64_t x = x0 + k;
            if (x >= a.nx) break;  // uniform
            const int32_t vi = a.vlut[cur[k]];
            uint32_t cnt;
            if (wave_group((uint32_t)vi, zok && vi >= 0, cnt)) {
                atomicAdd(&hx[k * a.vb + vi], cnt);
                const int64_t r = isqrt64((x - a.cx) * (x - a.cx) + dy2);
                if (r < a.Nr) atomicAdd(&rb[r * a.vb + vi], (unsigned long long)cnt);
            }
        }
#pragma unroll
        for (int k = 0; k < COLS; ++k) cur[k] = nxt[k];
    }
    __syncthreads();
    for (int64_t i = threadIdx.x; i < kXrCols * a.vb; i += kXrThreads) {
        const uint32_t c = hx[i];
        const int64_t x = x0 + i / a.vb;
        if (c != 0 && x < a.nx) atomicAdd(&xb[x * a.vb + i % a.vb], (unsigned long long)c);
    }
}

// Host-side bin tables: histograms.cc:61-62 (:278-279) and :292.
static void voxel_lut(double vmin, double vmax, int64_t vb, int32_t *lut) {
    for (int64_t i = 0; i < 65536; ++i) {
        const uint16_t raw = (uint16_t)i;
        const uint16_t voxel = (raw >= vmin && raw <= vmax) ? raw : (uint16_t)0;
        const int64_t idx = (int64_t)std::floor(static_cast<double>(vb - 1) * ((voxel - vmin) / (vmax - vmin)));
        lut[i] = (voxel != 0 && idx >= 0 && idx < vb) ? (int32_t)idx : -1;
    }
}

static void field_lut(double fmin, double fmax, int64_t fb, int32_t *lut) {
    for (int64_t i = 0; i < 65536; ++i) {
        const uint16_t f = (uint16_t)i;
        const int64_t idx = (int64_t)std::floor(static_cast<double>(fb - 1) * ((f - fmin) / (fmax - fmin)));
        lut[i] = (f > 0 && idx >= 0 && idx < fb) ? (int32_t)idx : -1;
    }
}

static int vec_ok(const void *p, int64_t nx) {
    return (nx % 8 == 0 && reinterpret_cast<uintptr_t>(p) % 16 == 0) ? 1 : 0;
}

static unsigned grid_of(const RowWalk &w) {
    return (unsigned)(w.rows * w.chunks);
}

// Reductions: a grid-stride grid of at most 8192 blocks (one atomic each).
static unsigned reduce_grid(const RowWalk &w) {
    return (unsigned)std::min<int64_t>(w.rows * w.chunks, 8192);
}

#define MXB_FHIP(expr)                                                          \
    do {                                                                        \
        hipError_t _e = (expr);                                                 \
        if (_e != hipSuccess) {                                                 \
            set_error(std::string(#expr) + ": " + hipGetErrorString(_e));       \
            return _e == hipErrorOutOfMemory ? ENOMEM_ : EDEVICE;               \
        }                                                                       \
    } while (0)

// As MXB_FHIP, for calls made after a copy from a host table was queued on
// `s`: an error path first waits for the stream, so the copy never reads the
// table after it is freed.
#define MXB_FHIP_AFTER_COPY(expr, s)                                            \
    do {                                                                        \
        hipError_t _e = (expr);                                                 \
        if (_e != hipSuccess) {                                                 \
            set_error(std::string(#expr) + ": " + hipGetErrorString(_e));       \
            (void)hipStreamSynchronize(s);                                      \
            return _e == hipErrorOutOfMemory ? ENOMEM_ : EDEVICE;               \
        }                                                                       \
    } while (0)

static int check_grid(const RowWalk &w) {
    if (w.rows * w.chunks > 0x7fffffffLL) {
        set_error("volume too large for one launch of the row kernels");
        return EINVAL_;
    }
    return OK;
}

// The device combine, synchronising on `s` between passes (the pass
// parameters are computed on the host from the reductions).
static int device_combine(const uint16_t *g, const uint16_t *e, int64_t nz, int64_t ny, int64_t nx,
                          const uint8_t *cyl_dev, const uint8_t *cyl_host, uint16_t *out,
                          uint32_t *ws, hipStream_t s) {
    const int64_t n = nz * ny * nx;
    if (n == 0) return OK;
    const RowWalk w = row_walk(nz * ny, nx);
    if (int rc = check_grid(w)) return rc;
    const int vec = vec_ok(g, nx) & vec_ok(e, nx) & vec_ok(out, nx);
    hipLaunchKernelGGL(k_fields_init, dim3(1), dim3(64), 0, s, ws);
    hipLaunchKernelGGL(k_max2, dim3(reduce_grid(w)), dim3(kRowThreads), 0, s, g, e, w, vec, ws);
    MXB_FHIP(hipGetLastError());
    uint32_t h[6];
    MXB_FHIP(hipMemcpyAsync(h, ws, 8, hipMemcpyDeviceToHost, s));
    MXB_FHIP(hipStreamSynchronize(s));
    if (h[0] == 0 || h[1] == 0) {
        set_error("combine_fields: a field's maximum is zero");
        return EINVAL_;
    }
    const float gm = (float)h[0], em = (float)h[1];
    hipLaunchKernelGGL(k_combine_minmax, dim3(reduce_grid(w)), dim3(kRowThreads), 0, s, g, e, cyl_dev, w, ny,
                       vec, gm, em, ws);
    MXB_FHIP(hipGetLastError());
    MXB_FHIP(hipMemcpyAsync(h, ws, sizeof(h), hipMemcpyDeviceToHost, s));
    MXB_FHIP(hipStreamSynchronize(s));
    // inside / outside the cylinder (0975:39, a 2D mask broadcast over z)
    bool any_in = true, any_out = false;
    if (cyl_host != nullptr) {
        any_in = false;
        for (int64_t i = 0; i < ny * nx; ++i) {
            if (cyl_host[i] != 0) any_in = true;
            else any_out = true;
        }
    }
    CombineParams p{};
    p.gm = gm;
    p.em = em;
    const float cmin = ord2f(h[2]), cmax = ord2f(h[3]);
    p.a = std::fabs(cmin);       // combined += np.abs(combined.min())
    p.M = cmax + p.a;            // combined /= combined.max()
    float vmin = 0.0f, vmax = 0.0f;
    if (any_in) {
        const float wmin = (ord2f(h[4]) + p.a) / p.M;
        const float wmax = (ord2f(h[5]) + p.a) / p.M;
        vmin = any_out ? std::fmin(0.0f, wmin) : wmin;
        vmax = any_out ? std::fmax(0.0f, wmax) : wmax;
    }
    // min_max starts vmax at numeric_limits<float>::min(), the smallest
    // positive normal (general.hh:78)
    vmax = std::fmax(vmax, std::numeric_limits<float>::min());
    p.vmin = vmin;
    p.range = vmax - vmin;       // general.hh:100, (float)vmax - (float)vmin
    hipLaunchKernelGGL(k_combine_out, dim3(grid_of(w)), dim3(kRowThreads), 0, s, g, e, cyl_dev, w, ny, vec,
                       p, out);
    MXB_FHIP(hipGetLastError());
    return OK;
}

static int64_t field_ws_bytes(int64_t nZ, int64_t nY, int64_t nX) {
    return 2 * 65536 * 4 + (nZ + nY + nX) * 4 + 256;
}

static int device_field_histogram(const uint16_t *vox, int64_t nZ, int64_t nY, int64_t nX,
                                  const uint16_t *field, int64_t nz, int64_t ny, int64_t nx,
                                  uint64_t *bins, int64_t fb, int64_t vb, double vmin, double vmax,
                                  double fmin, double fmax, void *ws, hipStream_t s) {
    if (nZ * nY * nX == 0) return OK;
    if (nz < 1 || ny < 1 || nx < 1 || fb < 1 || vb < 1 || fb * vb >= (int64_t)0xffffffffLL) {
        set_error("field_histogram: invalid field shape or bin counts");
        return EINVAL_;
    }
    std::vector<int32_t> host(2 * 65536 + nZ + nY + nX);
    voxel_lut(vmin, vmax, vb, host.data());
    field_lut(fmin, fmax, fb, host.data() + 65536);
    // histograms.cc:246-249, :284-287
    const double dz = (double)nz / (double)nZ, dy = (double)ny / (double)nY, dx = (double)nx / (double)nX;
    int32_t *zm = host.data() + 2 * 65536, *ym = zm + nZ, *xm = ym + nY;
    for (int64_t Z = 0; Z < nZ; ++Z) zm[Z] = (int32_t)std::min<int64_t>(nz - 1, (int64_t)std::floor((double)Z * dz));
    for (int64_t Y = 0; Y < nY; ++Y) ym[Y] = (int32_t)std::min<int64_t>(ny - 1, (int64_t)std::floor((double)Y * dy));
    for (int64_t X = 0; X < nX; ++X) xm[X] = (int32_t)std::min<int64_t>(nx - 1, (int64_t)std::floor((double)X * dx));
    int32_t *d = (int32_t *)ws;
    MXB_FHIP(hipMemcpyAsync(d, host.data(), host.size() * 4, hipMemcpyHostToDevice, s));
    FieldHistArgs a{nZ, nY, nX, nz, ny, nx, vb, fb, d, d + 65536, d + 2 * 65536, d + 2 * 65536 + nZ,
                    d + 2 * 65536 + nZ + nY};
    const int64_t bricks = ((nX + kBrickX - 1) / kBrickX) * ((nY + kBrickY - 1) / kBrickY) *
                           ((nZ + kBrickZ - 1) / kBrickZ);
    const unsigned grid = (unsigned)std::min<int64_t>(bricks, kFieldHistBlocks);
    hipLaunchKernelGGL(k_field_histogram, dim3(grid), dim3(kRowThreads), 0, s, vox, field, a, vec_ok(vox, nX),
                       (unsigned long long *)bins);
    MXB_FHIP_AFTER_COPY(hipGetLastError(), s);
    // the host table must outlive the copy
    MXB_FHIP(hipStreamSynchronize(s));
    return OK;
}

static int device_axis_histograms(const uint16_t *vox, int64_t nz, int64_t ny, int64_t nx,
                                  int64_t z_offset, uint64_t *xb, uint64_t *yb, uint64_t *zb,
                                  uint64_t *rb, int64_t Nz, int64_t Nr, int64_t vb, uint64_t cy,
                                  uint64_t cx, double vmin, double vmax, void *ws, hipStream_t s) {
    if (nz * ny * nx == 0) return OK;
    if (vb < 1 || z_offset < 0 || Nz < 0 || Nr < 0) {
        set_error("axis_histograms: invalid arguments");
        return EINVAL_;
    }
    // every argument check precedes the copy of the lookup table
    const RowWalk w = row_walk(nz * ny, nx);
    if (int rc = check_grid(w)) return rc;
    if (vb > 8192) {
        set_error("axis_histograms: at most 8192 voxel bins");
        return EINVAL_;
    }
    const int cols = 8 * vb * 4 <= kXrLds ? 8 : 4;  // vb <= 8192: 4 columns fit
    const int64_t blocks = ((nx + cols - 1) / cols) * ((nz + kXrThreads - 1) / kXrThreads);
    if (blocks > 0x7fffffffLL) {
        set_error("axis_histograms: volume too large for one launch");
        return EINVAL_;
    }
    const int smem = (int)(cols * vb * 4);
    const void *fn = cols == 8 ? reinterpret_cast<const void *>(&k_axis_xr<8>)
                               : reinterpret_cast<const void *>(&k_axis_xr<4>);
    MXB_FHIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, smem));
    std::vector<int32_t> host(65536);
    voxel_lut(vmin, vmax, vb, host.data());
    int32_t *d = (int32_t *)ws;
    MXB_FHIP(hipMemcpyAsync(d, host.data(), host.size() * 4, hipMemcpyHostToDevice, s));
    AxisHistArgs a{nz, ny, nx, z_offset, Nz, Nr, vb, d, (int64_t)cy, (int64_t)cx};
    const unsigned grid = (unsigned)std::min<int64_t>(w.rows, 4096);
    hipLaunchKernelGGL(k_axis_histograms, dim3(grid), dim3(kRowThreads), (unsigned)(2 * vb * 4), s, vox, w, a,
                       vec_ok(vox, nx), (unsigned long long *)yb, (unsigned long long *)zb);
    MXB_FHIP_AFTER_COPY(hipGetLastError(), s);
    if (cols == 8)
        hipLaunchKernelGGL(k_axis_xr<8>, dim3((unsigned)blocks), dim3(kXrThreads), (unsigned)smem, s, vox, a,
                           vec_ok(vox, nx), (unsigned long long *)xb, (unsigned long long *)rb);
    else
        hipLaunchKernelGGL(k_axis_xr<4>, dim3((unsigned)blocks), dim3(kXrThreads), (unsigned)smem, s, vox, a,
                           vec_ok(vox, nx), (unsigned long long *)xb, (unsigned long long *)rb);
    MXB_FHIP_AFTER_COPY(hipGetLastError(), s);
    MXB_FHIP(hipStreamSynchronize(s));
    return OK;
}

// RAII device buffer for the host entries.
struct FBuf {
    void *p = nullptr;
    ~FBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t n) {
        if (n == 0) n = 1;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
            return e == hipErrorOutOfMemory ? ENOMEM_ : EDEVICE;
        }
        return OK;
    }
};

}  // namespace mxb

using namespace mxb;

extern "C" {

int64_t mxb_combine_workspace_bytes(void) { return 256; }

int mxb_device_combine_fields(const uint16_t *dev_gauss, const uint16_t *dev_edt, int64_t nz,
                              int64_t ny, int64_t nx, const uint8_t *dev_cylinder,
                              uint16_t *dev_out, void *dev_workspace, void *stream) {
    if (nz < 0 || ny < 0 || nx < 0 || (nz * ny * nx > 0 && (!dev_gauss || !dev_edt || !dev_out || !dev_workspace))) {
        set_error("combine_fields: invalid arguments");
        return EINVAL_;
    }
    std::vector<uint8_t> cyl;
    if (dev_cylinder != nullptr && ny * nx > 0) {
        // ordered after the caller's work on its stream (which may produce the mask)
        cyl.resize((size_t)(ny * nx));
        MXB_FHIP(hipMemcpyAsync(cyl.data(), dev_cylinder, cyl.size(), hipMemcpyDeviceToHost, (hipStream_t)stream));
        MXB_FHIP(hipStreamSynchronize((hipStream_t)stream));
    }
    return device_combine(dev_gauss, dev_edt, nz, ny, nx, dev_cylinder,
                          dev_cylinder != nullptr ? cyl.data() : nullptr, dev_out, (uint32_t *)dev_workspace,
                          (hipStream_t)stream);
}

int mxb_combine_fields(const uint16_t *gauss, const uint16_t *edt, int64_t nz, int64_t ny,
                       int64_t nx, const uint8_t *cylinder, uint16_t *out, int verbose) {
    (void)verbose;
    if (nz < 0 || ny < 0 || nx < 0 || (nz * ny * nx > 0 && (!gauss || !edt || !out))) {
        set_error("combine_fields: invalid arguments");
        return EINVAL_;
    }
    const int64_t n = nz * ny * nx;
    if (n == 0) return OK;
    FBuf g, e, o, c, ws;
    int rc;
    if ((rc = g.alloc(n * 2)) || (rc = e.alloc(n * 2)) || (rc = o.alloc(n * 2)) || (rc = ws.alloc(256))) return rc;
    if (cylinder != nullptr && (rc = c.alloc(ny * nx))) return rc;
    hipStream_t s = nullptr;
    MXB_FHIP(staged_copy(g.p, gauss, n * 2, true, nullptr));
    MXB_FHIP(staged_copy(e.p, edt, n * 2, true, nullptr));
    if (cylinder != nullptr) MXB_FHIP(hipMemcpy(c.p, cylinder, ny * nx, hipMemcpyHostToDevice));
    rc = device_combine((const uint16_t *)g.p, (const uint16_t *)e.p, nz, ny, nx,
                        cylinder != nullptr ? (const uint8_t *)c.p : nullptr, cylinder, (uint16_t *)o.p,
                        (uint32_t *)ws.p, s);
    if (rc != OK) return rc;
    MXB_FHIP(staged_copy(out, o.p, n * 2, false, nullptr));
    return OK;
}

int64_t mxb_histogram_workspace_bytes(int64_t nZ, int64_t nY, int64_t nX) {
    if (nZ < 0 || nY < 0 || nX < 0) return -1;
    return field_ws_bytes(nZ, nY, nX);
}

int mxb_device_field_histogram(const uint16_t *dev_voxels, int64_t nZ, int64_t nY, int64_t nX,
                               const uint16_t *dev_field, int64_t nz, int64_t ny, int64_t nx,
                               uint64_t *dev_bins, int64_t field_bins, int64_t voxel_bins,
                               double vmin, double vmax, double fmin, double fmax,
                               void *dev_workspace, void *stream) {
    if (nZ < 0 || nY < 0 || nX < 0 || (nZ * nY * nX > 0 && (!dev_voxels || !dev_field || !dev_bins || !dev_workspace))) {
        set_error("field_histogram: invalid arguments");
        return EINVAL_;
    }
    return device_field_histogram(dev_voxels, nZ, nY, nX, dev_field, nz, ny, nx, dev_bins, field_bins,
                                  voxel_bins, vmin, vmax, fmin, fmax, dev_workspace, (hipStream_t)stream);
}

int mxb_field_histogram(const uint16_t *voxels, int64_t nZ, int64_t nY, int64_t nX,
                        const uint16_t *field, int64_t nz, int64_t ny, int64_t nx, uint64_t *bins,
                        int64_t field_bins, int64_t voxel_bins, double vmin, double vmax,
                        double fmin, double fmax, int verbose) {
    (void)verbose;
    if (nZ < 0 || nY < 0 || nX < 0 || nz < 0 || ny < 0 || nx < 0 || field_bins < 1 || voxel_bins < 1 ||
        (nZ * nY * nX > 0 && (!voxels || !field || !bins))) {
        set_error("field_histogram: invalid arguments");
        return EINVAL_;
    }
    const int64_t n = nZ * nY * nX, nf = nz * ny * nx, nb = field_bins * voxel_bins;
    if (n == 0) return OK;
    FBuf v, f, b, ws;
    int rc;
    if ((rc = v.alloc(n * 2)) || (rc = f.alloc(nf * 2)) || (rc = b.alloc(nb * 8)) ||
        (rc = ws.alloc(field_ws_bytes(nZ, nY, nX))))
        return rc;
    MXB_FHIP(staged_copy(v.p, voxels, n * 2, true, nullptr));
    MXB_FHIP(staged_copy(f.p, field, nf * 2, true, nullptr));
    MXB_FHIP(hipMemcpy(b.p, bins, nb * 8, hipMemcpyHostToDevice));
    rc = device_field_histogram((const uint16_t *)v.p, nZ, nY, nX, (const uint16_t *)f.p, nz, ny, nx,
                                (uint64_t *)b.p, field_bins, voxel_bins, vmin, vmax, fmin, fmax, ws.p, nullptr);
    if (rc != OK) return rc;
    MXB_FHIP(hipMemcpy(bins, b.p, nb * 8, hipMemcpyDeviceToHost));
    return OK;
}

int mxb_device_axis_histograms(const uint16_t *dev_voxels, int64_t nz, int64_t ny, int64_t nx,
                               int64_t z_offset, uint64_t *dev_x_bins, uint64_t *dev_y_bins,
                               uint64_t *dev_z_bins, uint64_t *dev_r_bins, int64_t Nz, int64_t Nr,
                               int64_t voxel_bins, uint64_t cy, uint64_t cx, double vmin,
                               double vmax, void *dev_workspace, void *stream) {
    if (nz < 0 || ny < 0 || nx < 0 ||
        (nz * ny * nx > 0 && (!dev_voxels || !dev_x_bins || !dev_y_bins || !dev_z_bins || !dev_r_bins || !dev_workspace))) {
        set_error("axis_histograms: invalid arguments");
        return EINVAL_;
    }
    return device_axis_histograms(dev_voxels, nz, ny, nx, z_offset, dev_x_bins, dev_y_bins, dev_z_bins,
                                  dev_r_bins, Nz, Nr, voxel_bins, cy, cx, vmin, vmax, dev_workspace,
                                  (hipStream_t)stream);
}

int mxb_axis_histograms(const uint16_t *voxels, int64_t nz, int64_t ny, int64_t nx,
                        int64_t z_offset, uint64_t *x_bins, uint64_t *y_bins, uint64_t *z_bins,
                        uint64_t *r_bins, int64_t Nz, int64_t Nr, int64_t voxel_bins, uint64_t cy,
                        uint64_t cx, double vmin, double vmax, int verbose) {
    (void)verbose;
    if (nz < 0 || ny < 0 || nx < 0 || voxel_bins < 1 || Nz < 0 || Nr < 0 ||
        (nz * ny * nx > 0 && (!voxels || !x_bins || !y_bins || !z_bins || !r_bins))) {
        set_error("axis_histograms: invalid arguments");
        return EINVAL_;
    }
    const int64_t n = nz * ny * nx;
    if (n == 0) return OK;
    FBuf v, xb, yb, zb, rb, ws;
    int rc;
    if ((rc = v.alloc(n * 2)) || (rc = xb.alloc(nx * voxel_bins * 8)) || (rc = yb.alloc(ny * voxel_bins * 8)) ||
        (rc = zb.alloc(Nz * voxel_bins * 8)) || (rc = rb.alloc(Nr * voxel_bins * 8)) ||
        (rc = ws.alloc(65536 * 4)))
        return rc;
    MXB_FHIP(staged_copy(v.p, voxels, n * 2, true, nullptr));
    MXB_FHIP(hipMemcpy(xb.p, x_bins, nx * voxel_bins * 8, hipMemcpyHostToDevice));
    MXB_FHIP(hipMemcpy(yb.p, y_bins, ny * voxel_bins * 8, hipMemcpyHostToDevice));
    if (Nz > 0) MXB_FHIP(hipMemcpy(zb.p, z_bins, Nz * voxel_bins * 8, hipMemcpyHostToDevice));
    if (Nr > 0) MXB_FHIP(hipMemcpy(rb.p, r_bins, Nr * voxel_bins * 8, hipMemcpyHostToDevice));
    rc = device_axis_histograms((const uint16_t *)v.p, nz, ny, nx, z_offset, (uint64_t *)xb.p, (uint64_t *)yb.p,
                                (uint64_t *)zb.p, (uint64_t *)rb.p, Nz, Nr, voxel_bins, cy, cx, vmin, vmax,
                                ws.p, nullptr);
    if (rc != OK) return rc;
    MXB_FHIP(hipMemcpy(x_bins, xb.p, nx * voxel_bins * 8, hipMemcpyDeviceToHost));
    MXB_FHIP(hipMemcpy(y_bins, yb.p, ny * voxel_bins * 8, hipMemcpyDeviceToHost));
    if (Nz > 0) MXB_FHIP(hipMemcpy(z_bins, zb.p, Nz * voxel_bins * 8, hipMemcpyDeviceToHost));
    if (Nr > 0) MXB_FHIP(hipMemcpy(r_bins, rb.p, Nr * voxel_bins * 8, hipMemcpyDeviceToHost));
    return OK;
}

}  // extern "C"
