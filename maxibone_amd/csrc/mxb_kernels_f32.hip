// FP32 line-pass kernels of the masked 3D diffusion (gfx950).
//
// One 1D correlation pass (reference: src/lib/cpp/cpu/diffusion.cc:31-60)
//   out[p] = sum_{r=-R..R} in[p + r] * k[R + r],   in[] = 0 outside the line,
// runs over "lines" whose first voxels sit at consecutive addresses (lane
// stride 1), so that a wave's load of one position covers contiguous bytes.
// A thread marches along its line over a segment of positions and keeps the
// inputs it still needs in registers: every input is fetched from memory once
// per segment.  The y-pass writes (z, x, y) order and the x-pass reads it back
// and writes (z, y, x), so that all three passes read along a contiguous axis.
//
// Arithmetic forms (all sum the taps r-ascending from 0, like cpu:46-54):
//   EXACT: product rounded, then added, no FMA: bit-identical to the reference
//          CPU build (-O3 without -march has no FMA contraction).
//   FMA:   one fused multiply-add per tap.
//   MFMA:  banded-Toeplitz tiles on v_mfma_f32_16x16x4_f32.  A wave owns 16
//          lines; per tile of 16 consecutive outputs of each line it runs
//          C = (16 + 2R)/4 MFMAs D[i][j] += T_c[i][kk] * X_c[kk][j] with
//          T_c[i][kk] = k[4c + kk - i] (zero outside the band) and
//          X_c[kk][j] = in_j(base - R + 4c + kk).
//
// Zero boundary: positions outside [0, n_ax) contribute 0*k, which adds +0
// and leaves every partial sum unchanged, so zero-padding equals the
// reference's skipped taps exactly.  The compiled tap capacity RT may exceed
// the runtime radius R: the taps are centred and padded with zeros, which
// again only adds exact zeros.
#include "mxb_internal.h"

namespace mxb {

typedef float v4f __attribute__((ext_vector_type(4)));

template <int N>
struct Taps {
    float k[N];
};

__device__ __forceinline__ float mac_fma(float acc, float w, float k) {
    return __builtin_fmaf(w, k, acc);
}

__device__ __forceinline__ float mac_exact(float acc, float w, float k) {
#pragma clang fp contract(off)
    const float p = w * k;
    return acc + p;
}

// floor(u * 65535.0f) as uint16 (cpu:191-196), converted through int32 so that
// out-of-range values wrap the way the x86 reference build does.
__device__ __forceinline__ uint16_t quantise_u16(float v) {
#pragma clang fp contract(off)
    const float s = v * 65535.0f;
    return (uint16_t)(int32_t)floorf(s);
}

__device__ __forceinline__ void epilogue_store(const Epilogue &e, int64_t off, float v) {
    if (e.mask != nullptr && e.mask[off] != 0) v = 1.0f;  // illuminate, cpu:295-302
    if (e.out_f32 != nullptr) e.out_f32[off] = v;
    if (e.out_u16 != nullptr) e.out_u16[off] = quantise_u16(v);
}

__device__ __forceinline__ int wave_id() {
    return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

// Block -> (segment, group of 64 lines, outer index) for the VALU form, where
// a block holds 4 waves of 64 lines each, one outer index per wave.
struct BlockPos {
    int64_t outer, lane0, a0, a1;
};

__device__ __forceinline__ BlockPos decode_valu(const LineGeom &g) {
    int64_t bid = blockIdx.x;
    const int64_t seg_id = bid % g.nseg;
    bid /= g.nseg;
    const int64_t nlb = (g.n_lane + 63) >> 6;
    const int64_t lb = bid % nlb;
    bid /= nlb;
    BlockPos p;
    p.outer = bid * 4 + wave_id();
    p.lane0 = lb * 64;
    p.a0 = seg_id * g.seg;
    p.a1 = p.a0 + g.seg < g.n_ax ? p.a0 + g.seg : g.n_ax;
    return p;
}

// ---------------------------------------------------------------- VALU form
template <int RT, int B, bool EXACT>
__global__ __launch_bounds__(256) void k_line_valu(const float *__restrict__ in, LineGeom g,
                                                   Epilogue e, Taps<2 * RT + 1> taps) {
    const BlockPos bp = decode_valu(g);
    if (bp.outer >= g.n_outer) return;  // wave-uniform
    const int lane = threadIdx.x & 63;
    const int64_t li = bp.lane0 + lane;
    if (li >= g.n_lane) return;  // no cross-lane work in this form
    constexpr int W = B + 2 * RT;
    const int64_t n_ax = g.n_ax, in_ax = g.in_ax;
    const float *src = in + bp.outer * g.in_outer + bp.lane0;  // wave-uniform base
    const int64_t obase = bp.outer * g.out_outer + li * g.out_lane;
    float w[W];
#pragma unroll
    for (int i = 0; i < 2 * RT; ++i) {
        const int64_t p = bp.a0 - RT + i;
        w[i] = (p >= 0 && p < n_ax) ? src[p * in_ax + lane] : 0.0f;
    }
    float nb[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int64_t p = bp.a0 + RT + b;
        nb[b] = p < n_ax ? src[p * in_ax + lane] : 0.0f;
    }
    for (int64_t a = bp.a0; a < bp.a1; a += B) {
#pragma unroll
        for (int b = 0; b < B; ++b) w[2 * RT + b] = nb[b];
        // prefetch the inputs of the next chunk
        if (a + B + RT + B <= n_ax) {
#pragma unroll
            for (int b = 0; b < B; ++b) nb[b] = src[(a + B + RT + b) * in_ax + lane];
        } else {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int64_t p = a + B + RT + b;
                nb[b] = p < n_ax ? src[p * in_ax + lane] : 0.0f;
            }
        }
        float acc[B];
#pragma unroll
        for (int b = 0; b < B; ++b) acc[b] = 0.0f;
#pragma unroll
        for (int t = 0; t <= 2 * RT; ++t) {
            const float kt = taps.k[t];
#pragma unroll
            for (int b = 0; b < B; ++b)
                acc[b] = EXACT ? mac_exact(acc[b], w[b + t], kt) : mac_fma(acc[b], w[b + t], kt);
        }
#pragma unroll
        for (int b = 0; b < B; ++b)
            if (a + b < bp.a1) epilogue_store(e, obase + (a + b) * g.out_ax, acc[b]);
#pragma unroll
        for (int i = 0; i < 2 * RT; ++i) w[i] = w[i + B];
    }
}

// ---------------------------------------------------------------- MFMA form
// Block = 4 waves = 64 consecutive lines at one outer index; wave w owns lines
// col0 = lane0 + 16w .. +15.  Lane l: column j = l % 16, row group kg = l / 16.
//   A operand (taps):   lane holds T_c[i = l%16][kk = l/16] = k[4c + l/16 - l%16]
//   B operand (inputs): lane holds X_c[kk = l/16][j = l%16]
//   D (outputs):        lane holds out_j(base + 4*(l/16) + r), r = 0..3
// The inputs of chunk G (positions a0 - R + 4G + kk) live in ring slot G % CR.
// NT tiles are processed together (independent accumulators, so the MFMA
// pipeline never waits on one accumulation chain), and the ring holds one
// group of prefetch beyond the chunks a group consumes.
struct BlockPosM {
    int64_t outer, col0, a0, a1;
};

__device__ __forceinline__ BlockPosM decode_mfma(const LineGeom &g) {
    int64_t bid = blockIdx.x;
    const int64_t seg_id = bid % g.nseg;
    bid /= g.nseg;
    const int64_t nlb = (g.n_lane + 63) >> 6;
    const int64_t lb = bid % nlb;
    bid /= nlb;
    BlockPosM p;
    p.outer = bid;
    p.col0 = lb * 64 + 16 * wave_id();
    p.a0 = seg_id * g.seg;
    p.a1 = p.a0 + g.seg < g.n_ax ? p.a0 + g.seg : g.n_ax;
    return p;
}

// 64 zero floats (device globals are zero-initialised).
__device__ float g_zero_line[64];

template <int RT>
struct MfmaShape {
    static constexpr int C = (16 + 2 * RT + 3) / 4;              // chunks per tile
    static constexpr int NT = 2;                                  // tiles per group
    static constexpr int GC = 4 * NT;                             // new chunks per group
    static constexpr int USE = C + 4 * (NT - 1);                  // chunks a group reads
    static constexpr int CR = ((USE + GC - 1) / GC) * GC + GC;    // ring slots
    static constexpr int P = CR / GC;                             // groups per ring period
};

__device__ __forceinline__ float select_mask(uint32_t m, int byte, float v) {
    return ((m >> (8 * byte)) & 0xffu) != 0 ? 1.0f : v;
}

template <int RT>
__global__ __launch_bounds__(256) void k_line_mfma(const float *__restrict__ in, LineGeom g,
                                                   Epilogue e, Taps<2 * RT + 1> taps) {
    using S = MfmaShape<RT>;
    constexpr int C = S::C, NT = S::NT, GC = S::GC, USE = S::USE, CR = S::CR, P = S::P;
    __shared__ float tap_lds[4 * C + 32];
    const BlockPosM bp = decode_mfma(g);
    const int lane = threadIdx.x & 63;
    const int col = lane & 15, kg = lane >> 4;
    // Zero-padded tap table: tap_lds[t + 16] = k[t] for t in [0, 2RT].
    for (int i = threadIdx.x; i < 4 * C + 32; i += 256) {
        const int t = i - 16;
        tap_lds[i] = (t >= 0 && t <= 2 * RT) ? taps.k[t] : 0.0f;
    }
    __syncthreads();
    float tv[C];
#pragma unroll
    for (int c = 0; c < C; ++c) tv[c] = tap_lds[4 * c + kg - col + 16];

    const int64_t n_ax = g.n_ax;
    const int64_t in_ax_b = g.in_ax * 4;  // bytes between positions (uniform)
    const int64_t li = bp.col0 + col;
    const bool act = li < g.n_lane;
    // Inactive columns read a valid column; their outputs are never stored and
    // an MFMA column j only feeds output column j.
    const int64_t lc = act ? li : g.n_lane - 1;
    const char *wbase = reinterpret_cast<const char *>(in + bp.outer * g.in_outer);  // uniform
    const uint32_t lcol_b = (uint32_t)(lc * 4);
    const uint32_t loff = lcol_b + (uint32_t)(kg * in_ax_b);
    const int64_t q_first = bp.a0 - RT;  // position of chunk 0, row group 0

    float ring[CR];
    // Positions outside the line read from a zero-filled line instead: the
    // address is selected before the load, so no value waits on a select and
    // the prefetched loads stay in flight across the next group's MFMAs.
    const char *zsrc = reinterpret_cast<const char *>(g_zero_line) + 4 * lane;
    auto load_chunk = [&](int slot, int64_t G) {
        const int64_t q0 = q_first + 4 * G;  // uniform
        const char *cbase = wbase + q0 * in_ax_b;
        const char *addr = cbase + loff;
        if (!(q0 >= 0 && q0 + 3 < n_ax)) {
            const int64_t p = q0 + kg;
            addr = (p >= 0 && p < n_ax) ? addr : zsrc;
        }
        ring[slot] = *reinterpret_cast<const float *>(addr);
    };
#pragma unroll
    for (int c = 0; c < USE + GC; ++c) load_chunk(c % CR, c);

    // Output addressing: lane base = outer*out_outer + li*out_lane + (a0 + 4kg)*out_ax.
    const int64_t obase = bp.outer * g.out_outer + lc * g.out_lane + (bp.a0 + 4 * kg) * g.out_ax;
    const int64_t ntiles = (bp.a1 - bp.a0 + 15) / 16;
    const int64_t ngroups = (ntiles + NT - 1) / NT;
    int64_t gi = 0;
    while (true) {
#pragma unroll
        for (int j = 0; j < P; ++j) {
            v4f acc[NT];
#pragma unroll
            for (int u = 0; u < NT; ++u) acc[u] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int c = 0; c < C; ++c)
#pragma unroll
                for (int u = 0; u < NT; ++u)
                    acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(tv[c], ring[(GC * j + 4 * u + c) % CR],
                                                                  acc[u], 0, 0, 0);
            // prefetch the new chunks of group gi + 2
#pragma unroll
            for (int q = 0; q < GC; ++q) load_chunk((GC * j + USE + GC + q) % CR, GC * gi + USE + GC + q);
#pragma unroll
            for (int u = 0; u < NT; ++u) {
                const int64_t t = NT * gi + u;
                const int64_t tpos = bp.a0 + 16 * t;  // uniform
                if (tpos >= bp.a1) break;
                const bool full = tpos + 16 <= bp.a1;
                if (act && full && g.vec_out) {
                    const int64_t off = obase + 16 * t;
                    v4f v = acc[u];
                    if (e.mask != nullptr) {
                        const uint32_t m = *reinterpret_cast<const uint32_t *>(e.mask + off);
                        v.x = select_mask(m, 0, v.x);
                        v.y = select_mask(m, 1, v.y);
                        v.z = select_mask(m, 2, v.z);
                        v.w = select_mask(m, 3, v.w);
                    }
                    if (e.out_f32 != nullptr) *reinterpret_cast<v4f *>(e.out_f32 + off) = v;
                    if (e.out_u16 != nullptr) {
                        const uint32_t lo = (uint32_t)quantise_u16(v.x) | ((uint32_t)quantise_u16(v.y) << 16);
                        const uint32_t hi = (uint32_t)quantise_u16(v.z) | ((uint32_t)quantise_u16(v.w) << 16);
                        *reinterpret_cast<uint2 *>(e.out_u16 + off) = make_uint2(lo, hi);
                    }
                } else if (act) {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (tpos + 4 * kg + r < bp.a1)
                            epilogue_store(e, obase + (16 * t + r) * g.out_ax, acc[u][r]);
                }
            }
            if (++gi == ngroups) return;
        }
    }
}

// ------------------------------------------------------ per-voxel, any radius
template <bool EXACT>
__global__ __launch_bounds__(256) void k_line_generic(const float *__restrict__ in, LineGeom g,
                                                      Epilogue e, Taps<kMaxTapsGeneric> taps,
                                                      int R) {
    const BlockPos bp = decode_valu(g);
    if (bp.outer >= g.n_outer) return;
    const int64_t li = bp.lane0 + (threadIdx.x & 63);
    if (li >= g.n_lane) return;
    const float *src = in + bp.outer * g.in_outer + li;
    const int64_t obase = bp.outer * g.out_outer + li * g.out_lane;
    for (int64_t p = bp.a0; p < bp.a1; ++p) {
        const int64_t lo = p < R ? -p : -(int64_t)R;
        const int64_t hi = (g.n_ax - 1 - p) < R ? (g.n_ax - 1 - p) : (int64_t)R;
        float sum = 0.0f;
        for (int64_t r = lo; r <= hi; ++r) {
            const float v = src[(p + r) * g.in_ax];
            sum = EXACT ? mac_exact(sum, v, taps.k[R + r]) : mac_fma(sum, v, taps.k[R + r]);
        }
        epilogue_store(e, obase + p * g.out_ax, sum);
    }
}

// ------------------------------------------------------------ element-wise
// u = (m > 0) as float (cpu:243-248); 8 elements per thread (an 8-byte mask
// load, two 16-byte stores) when `vec` says the pointers allow it, then the
// tail (or everything) one element per thread.
__global__ __launch_bounds__(256) void k_init_state(const uint8_t *__restrict__ m,
                                                    float *__restrict__ u, int64_t n, int vec) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t done = 0;
    if (vec) {
        const int64_t n8 = n >> 3;
        for (int64_t g = tid; g < n8; g += stride) {
            const uint2 mv = *reinterpret_cast<const uint2 *>(m + 8 * g);
            float r[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) r[e] = (((e < 4 ? mv.x : mv.y) >> (8 * (e & 3))) & 0xffu) ? 1.0f : 0.0f;
            float4 *dst = reinterpret_cast<float4 *>(u + 8 * g);
            dst[0] = make_float4(r[0], r[1], r[2], r[3]);
            dst[1] = make_float4(r[4], r[5], r[6], r[7]);
        }
        done = n8 << 3;
    }
    for (int64_t i = done + tid; i < n; i += stride) u[i] = m[i] > 0 ? 1.0f : 0.0f;
}

__global__ __launch_bounds__(256) void k_quantise(const float *__restrict__ u,
                                                  uint16_t *__restrict__ q, float *__restrict__ f,
                                                  int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float v = u[i];
        if (q != nullptr) q[i] = quantise_u16(v);
        if (f != nullptr) f[i] = v;
    }
}

// The reference on-disk path's mask rule (store_mask, cpu:331-337): a voxel
// is illuminated when its state is exactly 1.0f.
__global__ __launch_bounds__(256) void k_mask_from_state(const float *__restrict__ u,
                                                         uint8_t *__restrict__ m, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        m[i] = u[i] == 1.0f ? 1 : 0;
}

// ---------------------------------------------------------------- launchers
static inline int64_t grid_stride_blocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    return b < 1 ? 1 : (b > 65536 ? 65536 : b);
}

hipError_t launch_init_state(const uint8_t *mask, float *u, int64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const int vec = reinterpret_cast<uintptr_t>(mask) % 8 == 0 && reinterpret_cast<uintptr_t>(u) % 16 == 0;
    hipLaunchKernelGGL(k_init_state, dim3((unsigned)grid_stride_blocks(vec ? (n + 7) / 8 : n)), dim3(256), 0, s,
                       mask, u, n, vec);
    return hipGetLastError();
}

hipError_t launch_mask_from_state(const float *u, uint8_t *m, int64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mask_from_state, dim3((unsigned)grid_stride_blocks(n)), dim3(256), 0, s, u, m, n);
    return hipGetLastError();
}

hipError_t launch_quantise(const float *u, uint16_t *q, float *f, int64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_quantise, dim3((unsigned)grid_stride_blocks(n)), dim3(256), 0, s, u, q, f, n);
    return hipGetLastError();
}

// Pack a mask into bits, words interleaved across rows (the MaskBits layout).
// Thread (y, w, z) builds word w of row (z, y) from its 32 bytes, so the
// threads of a block store consecutive words (consecutive y); each reads one
// whole 32-byte run.  Rows whose runs are 16-byte aligned take two 16-byte
// loads; the rest, and a row's partial last word, take bytes.

// Bit i of the result: byte i of v is nonzero (i = 0 .. 3).
__device__ __forceinline__ uint32_t nonzero_bytes4(uint32_t v) {
    const uint32_t t = (((v & 0x7f7f7f7fu) + 0x7f7f7f7fu) | v) & 0x80808080u;
    return ((t >> 7) & 1u) | ((t >> 14) & 2u) | ((t >> 21) & 4u) | ((t >> 28) & 8u);
}

__global__ __launch_bounds__(256) void k_pack_mask_bits(const uint8_t *__restrict__ m, int64_t nz, int64_t ny,
                                                        int64_t nx, int64_t mask_row, int64_t mask_plane,
                                                        uint32_t *__restrict__ bits) {
    const int64_t y = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t w = blockIdx.y;
    const int64_t wpr = (nx + 31) / 32;
    if (y >= ny) return;
    const int64_t x0 = 32 * w;
    for (int64_t z = blockIdx.z; z < nz; z += gridDim.z) {
        const uint8_t *row = m + z * mask_plane + y * mask_row + x0;
        uint32_t word = 0;
        if (x0 + 32 <= nx && (reinterpret_cast<uintptr_t>(row) & 15) == 0) {
            const uint4 a = *reinterpret_cast<const uint4 *>(row);
            const uint4 b = *reinterpret_cast<const uint4 *>(row + 16);
            const uint32_t d[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
            for (int k = 0; k < 8; ++k) word |= nonzero_bytes4(d[k]) << (4 * k);
        } else {
            const int n = (int)(nx - x0 < 32 ? nx - x0 : 32);
            for (int i = 0; i < n; ++i) word |= (row[i] != 0 ? 1u : 0u) << i;
        }
        bits[(z * wpr + w) * ny + y] = word;
    }
}

hipError_t launch_pack_mask_bits(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx, int64_t mask_row,
                                 int64_t mask_plane, uint32_t *bits, hipStream_t s) {
    if (nz * ny * nx == 0) return hipSuccess;
    const int64_t wpr = (nx + 31) / 32;
    if (wpr > 65535) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((ny + 255) / 256), (unsigned)wpr, (unsigned)(nz < 4096 ? nz : 4096));
    hipLaunchKernelGGL(k_pack_mask_bits, grid, dim3(256), 0, s, mask, nz, ny, nx, mask_row, mask_plane, bits);
    return hipGetLastError();
}

LineGeom geom_z(int64_t nz, int64_t ny, int64_t nx) {
    LineGeom g{};
    g.n_ax = nz; g.in_ax = ny * nx; g.out_lane = 1; g.out_ax = ny * nx;
    if (nx % 64 != 0 && (ny * nx) % 64 == 0) {
        // rows are not whole 64-lane groups but planes are: the z-pass reads
        // no mask and every (y, x) lane is an independent line, so a plane is
        // one run of ny * nx lanes, and each group is a whole aligned piece
        // instead of every row ending in a partial group whose cache lines
        // the next group reads and writes again
        g.n_lane = ny * nx; g.n_outer = 1;
        g.in_outer = ny * nx; g.out_outer = ny * nx;
    } else {
        g.n_lane = nx; g.n_outer = ny;
        g.in_outer = nx; g.out_outer = nx;
    }
    return g;
}

LineGeom geom_y(int64_t nz, int64_t ny, int64_t nx) {
    LineGeom g{};
    g.n_lane = nx; g.n_outer = nz; g.n_ax = ny;
    g.in_outer = ny * nx; g.in_ax = nx;
    g.out_lane = ny; g.out_outer = nx * ny; g.out_ax = 1;  // (z, x, y) order
    return g;
}

LineGeom geom_x(int64_t nz, int64_t ny, int64_t nx) {
    LineGeom g{};
    g.n_lane = ny; g.n_outer = nz; g.n_ax = nx;
    g.in_outer = nx * ny; g.in_ax = ny;                    // reads (z, x, y) order
    g.out_lane = nx; g.out_outer = ny * nx; g.out_ax = 1;  // writes (z, y, x)
    return g;
}

// Split lines into segments (multiples of 16 positions) until the launch holds
// enough threads to fill the chip: 256 CUs x `waves_per_cu` waves.  Each extra
// segment re-reads 2R halo positions.
static void plan_segments(LineGeom &g, int64_t lines_per_thread, int waves_per_cu) {
    const int64_t threads = ((g.n_lane + 63) / 64) * 64 * g.n_outer / lines_per_thread;
    const int64_t target = 256LL * waves_per_cu * 64;
    int64_t nseg = threads > 0 ? (target + threads - 1) / threads : 1;
    const int64_t max_seg = (g.n_ax + 63) / 64;
    if (nseg > max_seg) nseg = max_seg;
    if (nseg < 1) nseg = 1;
    int64_t seg = (g.n_ax + nseg - 1) / nseg;
    seg = ((seg + 15) / 16) * 16;
    if (seg < 16) seg = 16;
    g.seg = seg;
    g.nseg = (g.n_ax + seg - 1) / seg;
    if (g.nseg < 1) g.nseg = 1;
}

// Vector epilogue: 4 consecutive outputs per lane are contiguous and aligned.
static int32_t vec_ok(const LineGeom &g, const Epilogue &e) {
    if (g.out_ax != 1 || g.out_lane % 4 != 0 || g.out_outer % 4 != 0) return 0;
    if (e.out_f32 && (reinterpret_cast<uintptr_t>(e.out_f32) % 16) != 0) return 0;
    if (e.out_u16 && (reinterpret_cast<uintptr_t>(e.out_u16) % 8) != 0) return 0;
    if (e.mask && (reinterpret_cast<uintptr_t>(e.mask) % 4) != 0) return 0;
    return 1;
}

// The MFMA form addresses a lane's inputs with a 32-bit byte offset from a
// wave-uniform base: 3 positions plus one line group must fit.
static bool mfma_offsets_fit(const LineGeom &g) {
    return (g.in_ax * 3 + g.n_lane) * 4 < (int64_t(1) << 32);
}

template <int RT>
static Taps<2 * RT + 1> centred_taps(const float *k, int R) {
    Taps<2 * RT + 1> t;
    for (int i = 0; i < 2 * RT + 1; ++i) {
        const int r = i - RT;
        t.k[i] = (r >= -R && r <= R) ? k[R + r] : 0.0f;
    }
    return t;
}

template <int RT>
static hipError_t launch_rt(const float *in, LineGeom g, const Epilogue &e, const float *k, int R,
                            Arith arith, hipStream_t s) {
    const Taps<2 * RT + 1> t = centred_taps<RT>(k, R);
    const dim3 block(256);
    if (arith == Arith::MFMA && !mfma_offsets_fit(g)) arith = Arith::FMA;
    if (arith == Arith::MFMA) {
        plan_segments(g, 1, 16);
        g.vec_out = vec_ok(g, e);
        // one block per (segment, 64 lines, outer)
        const int64_t blocks = g.nseg * ((g.n_lane + 63) / 64) * g.n_outer;
        if (blocks == 0) return hipSuccess;
        hipLaunchKernelGGL((k_line_mfma<RT>), dim3((unsigned)blocks), block, 0, s, in, g, e, t);
        return hipGetLastError();
    }
    constexpr int B = RT <= 40 ? 16 : 8;
    plan_segments(g, 1, 8);
    const int64_t blocks = g.nseg * ((g.n_lane + 63) / 64) * ((g.n_outer + 3) / 4);
    if (blocks == 0) return hipSuccess;
    if (arith == Arith::EXACT)
        hipLaunchKernelGGL((k_line_valu<RT, B, true>), dim3((unsigned)blocks), block, 0, s, in, g, e, t);
    else
        hipLaunchKernelGGL((k_line_valu<RT, B, false>), dim3((unsigned)blocks), block, 0, s, in, g, e, t);
    return hipGetLastError();
}

static hipError_t launch_generic(const float *in, LineGeom g, const Epilogue &e, const float *k,
                                 int R, Arith arith, hipStream_t s) {
    plan_segments(g, 1, 8);
    const int64_t blocks = g.nseg * ((g.n_lane + 63) / 64) * ((g.n_outer + 3) / 4);
    if (blocks == 0) return hipSuccess;
    Taps<kMaxTapsGeneric> t;
    for (int i = 0; i < kMaxTapsGeneric; ++i) t.k[i] = i < 2 * R + 1 ? k[i] : 0.0f;
    const dim3 grid((unsigned)blocks), block(256);
    if (arith == Arith::EXACT)
        hipLaunchKernelGGL((k_line_generic<true>), grid, block, 0, s, in, g, e, t, R);
    else
        hipLaunchKernelGGL((k_line_generic<false>), grid, block, 0, s, in, g, e, t, R);
    return hipGetLastError();
}

hipError_t launch_line_pass_f32(const float *in, const LineGeom &g, const Epilogue &e,
                                const float *taps, int R, Arith arith, hipStream_t s) {
    if (arith == Arith::MFMA && !force_ring_kernels() && lds_pass_supported(in, g, R))
        return launch_line_pass_f32_lds(in, g, e, taps, R, s);
#define MXB_RT_CASE(rt) \
    if (R <= rt) return launch_rt<rt>(in, g, e, taps, R, arith, s);
    MXB_RT_CASE(2)
    MXB_RT_CASE(4)
    MXB_RT_CASE(6)
    MXB_RT_CASE(8)
    MXB_RT_CASE(10)
    MXB_RT_CASE(12)
    MXB_RT_CASE(14)
    MXB_RT_CASE(16)
    MXB_RT_CASE(20)
    MXB_RT_CASE(24)
    MXB_RT_CASE(28)
    MXB_RT_CASE(32)
    MXB_RT_CASE(36)
    MXB_RT_CASE(40)
    MXB_RT_CASE(48)
    MXB_RT_CASE(56)
    MXB_RT_CASE(60)
    MXB_RT_CASE(64)
#undef MXB_RT_CASE
    return launch_generic(in, g, e, taps, R, arith, s);
}

}  // namespace mxb
