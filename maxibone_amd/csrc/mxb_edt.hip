// Euclidean distance transform of a binary volume (gfx950): the core of the
// implant EDT field (src/processing_steps/0950_generate_edt.py:44-66, which
// calls edt.edt of the external `edt` package).  C ABI in
// include/maxibone_edt.h.
//
// Squared distances separably, three passes over (z, y, x) volumes:
//   pass x: g1[i] = min over sites j of the row of (i - j)^2
//   pass y: g2[i] = min_j g1[j] + (i - j)^2 along y
//   pass z: g3[i] = min_j g2[j] + (i - j)^2 along z
// Pass x is a row scan (k_edt_rows: per 64-position chunk a ballot of the
// sites, then the nearest site before and after each chunk by prefix max /
// suffix min over the chunks).
//
// Passes y and z are the lower envelope of the parabolas g[j] + (i - j)^2.
// Since g[j] + (i - j)^2 = i^2 + (c_j - 2 i j) with c_j = g[j] + j^2, the
// minimum over j is attained at a vertex of the lower convex hull of the
// points (j, c_j), and the minimising vertex moves right as i grows.  With
// at most 4096 voxels per axis, positions are below 2^12 and c below 2^27:
// 32-bit values, with 64-bit cross products.  One wave per line (64 lanes,
// a band of L positions per lane, 64 L >= n):
//   A. each lane builds the lower hull of its band (monotone chain; the
//      band's values preloaded into registers, the stack in LDS);
//   B. six rounds merge adjacent groups of bands: the lower bridge of two
//      x-separated hulls is found by walking the left end left and the right
//      end right until neither moves (the bands' emptiness from one ballot
//      per round); the merged hull is a prefix of the left hull and a suffix
//      of the right one (each band keeps a [lo, hi) range);
//   C. the surviving vertices are compacted (wave prefix sum);
//   D. vertex m + 1 beats vertex m from an integer switch position on, and
//      the switch positions grow with m: a table holding at each position the
//      largest m + 1 switching there, then a prefix max over the positions,
//      gives every position its vertex.
// The results stay in registers until the line is answered, then overwrite
// the line in LDS.
//
// A block has one wave per line and holds its lines in an LDS tile, loaded
// and stored with coalesced accesses (for the strided passes, consecutive x
// per position); the line count is the largest that keeps a block within
// kBlockBudget bytes of LDS.  Tiles are dealt to the 8 XCDs in contiguous
// ranges, so neighbouring tiles, which share cache lines, meet in one L2.
// Positions in a tile line are padded by one word per 16 (p + p / 16), and
// the band stacks are stored entry-major (entry k of band b at k * 64 + b),
// so that lanes hit different banks.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <string>

#include "maxibone_diffusion.h"
#include "maxibone_edt.h"
#include "mxb_internal.h"

namespace mxb {

namespace {

constexpr uint32_t kNone = MXB_EDT_NONE;
constexpr int kMaxLines = 8;              // lines (= waves) per block
constexpr int kMetaInts = 3 * 64;         // per wave: band lo, hi and spare
constexpr int kBlockBudget = 80 * 1024;   // LDS bytes per block (two blocks per CU)

struct EdtPass {
    const void *in;        // uint8 (first pass) or uint32
    uint32_t *out;         // uint32, or float bits (last pass)
    int64_t n;             // positions per line
    int64_t pos_stride;    // elements between positions
    int64_t n_inner, inner_stride;
    int64_t n_outer, outer_stride;
    int lines;             // lines per block (power of two)
    int pitch;             // LDS words per tile line
    const uint8_t *cyl;    // last pass: cylinder mask indexed by the line, or null
    uint32_t *minmax;      // last pass: ordered-encoded min and max
};

__device__ __forceinline__ int tix(int p) { return p + (p >> 4); }

// LDS ordering between the lanes of one wave (each wave's LDS operations
// execute in order; this keeps the compiler from moving them across).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t f2ord_d(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord2f_d(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// cross((x0,c0), (x1,c1), (x2,c2)) <= 0: the middle point is on or above the
// segment between the outer two (x0 < x1 < x2).  Positions are below 2^12
// and c = g + x^2 below 3 * 4096^2 < 2^26 + 2^24, so the operands are 32-bit
// and each product (below 2^39) comes from one 32 x 32 -> 64 multiply.
__device__ __forceinline__ bool not_below(int x0, int c0, int x1, int c1, int x2, int c2) {
    return (int64_t)(x1 - x0) * (c2 - c0) - (int64_t)(c1 - c0) * (x2 - x0) <= 0;
}

// Highest / lowest set bit of a nonzero 64-bit mask.
__device__ __forceinline__ int top_bit(uint64_t m) { return 63 - __builtin_clzll(m); }
__device__ __forceinline__ int low_bit(uint64_t m) { return __builtin_ctzll(m); }
__device__ __forceinline__ uint64_t bits_below(int k) { return k <= 0 ? 0ull : (k >= 64 ? ~0ull : (1ull << k) - 1); }

// Position from which hull vertex m + 1 beats vertex m (they tie at no
// integer before it): the least i with c1 - c0 < 2 i (x1 - x0), clamped to
// [0, n].
__device__ __forceinline__ int switch_pos(int x0, int c0, int x1, int c1, int n) {
    const int num = c1 - c0;
    if (num < 0) return 0;
    const int t = (int)((uint32_t)num / (uint32_t)(2 * (x1 - x0))) + 1;
    return t > n ? n : t;
}

// Envelope pass along y or z: uint32 -> uint32, or (LAST) uint32 -> float
// distance x cylinder with the block's min and max folded into P.minmax.
// Lines per block that the LDS budget admits for a band length L (the host
// clamps to it): a line needs 4 (pitch + 64 L + kMetaInts) bytes, over 25 KB
// once L = 64 (n > 2048) and over 13 KB at L = 32.  The kernel's launch bound
// follows it, so the long-line instances may hold their per-lane arrays
// (64 values each at L = 64) in registers without spilling.
constexpr int edt_max_lines(int L) { return L >= 64 ? 2 : (L >= 32 ? 4 : kMaxLines); }

template <int L, bool LAST>
__global__ __launch_bounds__(edt_max_lines(L) * 64) void k_edt_pass(EdtPass P) {
    extern __shared__ __attribute__((aligned(16))) uint32_t smem[];
    const int kThreads = P.lines * 64;  // one wave per line
    const int n = (int)P.n;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    uint32_t *const tile = smem;
    // per wave: stk (64 L uint16: the band stacks, entry k of band b at
    // k * 64 + b, later the switch table, position p at tpos(p)), env (64 L
    // uint16); lanes touch consecutive entries, so no bank conflicts
    uint16_t *const stk = reinterpret_cast<uint16_t *>(smem + P.lines * P.pitch) + wv * 128 * L;
    uint16_t *const env = stk + 64 * L;
    int *const blo = reinterpret_cast<int *>(smem + P.lines * P.pitch + P.lines * 64 * L) + wv * kMetaInts;
    auto tpos = [](int p) { return (p % L) * 64 + p / L; };
    int *const bhi = blo + 64;

    // tile of this block (XCD-aware order: block b runs on XCD b % 8)
    const int64_t chunks = (P.n_inner + P.lines - 1) / P.lines;
    const int64_t B = chunks * P.n_outer;
    const int64_t b = blockIdx.x, q = B / 8, rem = B % 8, xcd = b % 8;
    const int64_t t_id = xcd * q + (xcd < rem ? xcd : rem) + b / 8;
    const int64_t outer = t_id / chunks, inner0 = (t_id % chunks) * P.lines;
    const int nl = (int)(P.n_inner - inner0 < P.lines ? P.n_inner - inner0 : P.lines);
    const int64_t base = outer * P.outer_stride + inner0 * P.inner_stride;
    const bool lines_contig = P.inner_stride == 1 && P.pos_stride != 1;
    int lshift = 0;
    while ((1 << lshift) < P.lines) ++lshift;

    // ---- load
    if (lines_contig) {
        const int total = P.lines * n;
        for (int t = tid; t < total; t += kThreads) {
            const int l = t & (P.lines - 1), p = t >> lshift;
            if (l >= nl) continue;
            const int64_t o = base + l + (int64_t)p * P.pos_stride;
            tile[l * P.pitch + tix(p)] = static_cast<const uint32_t *>(P.in)[o];
        }
    } else {
        for (int l = 0; l < nl; ++l)
            for (int p = tid; p < n; p += kThreads) {
                const int64_t o = base + (int64_t)l * P.inner_stride + (int64_t)p * P.pos_stride;
                tile[l * P.pitch + tix(p)] = static_cast<const uint32_t *>(P.in)[o];
            }
    }
    __syncthreads();

    uint32_t omin = 0xffffffffu, omax = 0u;
    const int b0 = lane * L;
    for (int l = wv; l < nl; l += P.lines) {
        uint32_t *const g = tile + l * P.pitch;
        float cyl = 1.0f;
        if (LAST && P.cyl != nullptr) cyl = P.cyl[inner0 + l] != 0 ? 1.0f : 0.0f;
        auto cval = [&](int x) -> int { return (int)g[tix(x)] + x * x; };

        // A. lower hull of the band's points (the band's values preloaded)
        uint32_t gv[L];
#pragma unroll
        for (int t = 0; t < L; ++t) gv[t] = b0 + t < n ? g[tix(b0 + t)] : kNone;
        int h = 0;
        int xa = 0, ca = 0, xb = 0, cb = 0;  // second and top of the stack
#pragma unroll
        for (int t = 0; t < L; ++t) {
            const int j = b0 + t;
            if (gv[t] != kNone) {
                const int cj = (int)gv[t] + j * j;
                while (h >= 2 && not_below(xa, ca, xb, cb, j, cj)) {
                    --h;
                    xb = xa;
                    cb = ca;
                    if (h >= 2) {
                        xa = stk[(h - 2) * 64 + lane];
                        ca = cval(xa);
                    }
                }
                stk[h * 64 + lane] = (uint16_t)j;
                if (h >= 1) {
                    xa = xb;
                    ca = cb;
                }
                xb = j;
                cb = cj;
                ++h;
            }
        }
        int my_lo = 0, my_hi = h;
        blo[lane] = 0;
        bhi[lane] = h;
        wave_sync();

        // B. merge groups of bands [l0, mid) and [mid, l0 + gs); the bands'
        // emptiness at the start of a round comes from one ballot
        for (int gs = 2; gs <= 64; gs <<= 1) {
            const uint64_t ne = __ballot(my_hi > my_lo);
            if ((lane & (gs - 1)) == 0) {
                const int l0 = lane, mid = lane + gs / 2, r = lane + gs;
                const uint64_t left = ne & bits_below(mid) & ~bits_below(l0);
                const uint64_t right = ne & bits_below(r) & ~bits_below(mid);
                if (left != 0 && right != 0) {
                    int ba = top_bit(left), bb = low_bit(right);
                    int ia = bhi[ba] - 1, ib = blo[bb];
                    int ax = stk[ia * 64 + ba], bx = stk[ib * 64 + bb];
                    int ac = cval(ax), bc = cval(bx);
                    int lo_a = blo[ba], hi_b = bhi[bb];  // range ends of the current bands
                    for (;;) {
                        bool moved = false;
                        for (;;) {  // left end moves left while it is not below (prev, b)
                            int pa = ba, pi = ia - 1, plo = lo_a;
                            if (pi < lo_a) {
                                const uint64_t before = left & bits_below(ba);
                                if (before == 0) break;
                                pa = top_bit(before);
                                plo = blo[pa];
                                pi = bhi[pa] - 1;
                            }
                            const int px = stk[pi * 64 + pa], pc = cval(px);
                            if (!not_below(px, pc, ax, ac, bx, bc)) break;
                            ba = pa; ia = pi; ax = px; ac = pc; lo_a = plo;
                            moved = true;
                        }
                        for (;;) {  // right end moves right while it is not below (a, next)
                            int na = bb, ni = ib + 1, nhi = hi_b;
                            if (ni >= hi_b) {
                                const uint64_t after = right & ~bits_below(bb + 1);
                                if (after == 0) break;
                                na = low_bit(after);
                                ni = blo[na];
                                nhi = bhi[na];
                            }
                            const int qx = stk[ni * 64 + na], qc = cval(qx);
                            if (!not_below(ax, ac, bx, bc, qx, qc)) break;
                            bb = na; ib = ni; bx = qx; bc = qc; hi_b = nhi;
                            moved = true;
                        }
                        if (!moved) break;
                    }
                    bhi[ba] = ia + 1;
                    for (int k = ba + 1; k < mid; ++k) bhi[k] = blo[k];
                    blo[bb] = ib;
                    for (int k = mid; k < bb; ++k) blo[k] = bhi[k];
                }
            }
            wave_sync();
            my_lo = blo[lane];
            my_hi = bhi[lane];
        }

        // C. compact the hull into env
        const int cnt = my_hi - my_lo;
        int incl = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl, d, 64);
            if (lane >= d) incl += v;
        }
        const int H = __shfl(incl, 63, 64);
        const int off = incl - cnt;
        for (int k = 0; k < cnt; ++k) env[off + k] = stk[(my_lo + k) * 64 + lane];
        wave_sync();

        // D. minimising vertex of every position: vertex m + 1 takes over from
        // m at switch_pos(m); the switch table (in the stack's place) holds at
        // each position the largest m + 1 switching there, and a prefix max
        // over positions gives each position its vertex
        uint32_t res[L];
        if (H == 0) {
#pragma unroll
            for (int t = 0; t < L; ++t) res[t] = kNone;
        } else {
#pragma unroll
            for (int t = 0; t < L; ++t) stk[t * 64 + lane] = 0;
            wave_sync();
            // edge m = 64 k + lane; edge m + 1 is the next lane's (lane 63:
            // lane 0's of the next chunk); a sentinel n + 1 follows the last edge
            auto edge_switch = [&](int m) -> int {
                if (m >= H - 1) return n + 1;
                const int x0 = env[m], x1 = env[m + 1];
                return switch_pos(x0, cval(x0), x1, cval(x1), n);
            };
            int s_cur = edge_switch(lane);
            for (int m0 = 0; m0 < H - 1; m0 += 64) {
                const int s_nxt = edge_switch(m0 + 64 + lane);
                const int s_dn = __shfl_down(s_cur, 1, 64);
                const int s_first = __shfl(s_nxt, 0, 64);
                const int s_next = lane < 63 ? s_dn : s_first;
                const int m = m0 + lane;
                if (m < H - 1 && s_next > s_cur && s_cur < n) stk[tpos(s_cur)] = (uint16_t)(m + 1);
                s_cur = s_nxt;
            }
            wave_sync();
            // the lane's own maximum first; the running maximum is read again
            // below rather than held, so only res[] stays live (L = 64 would
            // otherwise need 2 x 64 VGPRs and spill)
            int run = 0;
#pragma unroll
            for (int t = 0; t < L; ++t) run = max(run, (int)stk[t * 64 + lane]);
            int carry = run;  // exclusive prefix max over lanes
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int v = __shfl_up(carry, d, 64);
                if (lane >= d) carry = max(carry, v);
            }
            carry = __shfl_up(carry, 1, 64);
            if (lane == 0) carry = 0;
            run = carry;
#pragma unroll
            for (int t = 0; t < L; ++t) {
                const int i = b0 + t;
                run = max(run, (int)stk[t * 64 + lane]);
                const int xm = env[run];
                res[t] = g[tix(xm)] + (uint32_t)((i - xm) * (i - xm));
            }
        }
        wave_sync();
#pragma unroll
        for (int t = 0; t < L; ++t) {
            const int i = b0 + t;
            if (i < n) {
                if (LAST) {
                    const float d = res[t] == kNone ? INFINITY : sqrtf((float)res[t]);
                    const float f = d * cyl;
                    const uint32_t o = f2ord_d(f);
                    omin = min(omin, o);
                    omax = max(omax, o);
                    g[tix(i)] = __float_as_uint(f);
                } else {
                    g[tix(i)] = res[t];
                }
            }
        }
        wave_sync();
    }
    __syncthreads();

    // ---- store
    if (lines_contig) {
        const int total = P.lines * n;
        for (int t = tid; t < total; t += kThreads) {
            const int l = t & (P.lines - 1), p = t >> lshift;
            if (l >= nl) continue;
            P.out[base + l + (int64_t)p * P.pos_stride] = tile[l * P.pitch + tix(p)];
        }
    } else {
        for (int l = 0; l < nl; ++l)
            for (int p = tid; p < n; p += kThreads)
                P.out[base + (int64_t)l * P.inner_stride + (int64_t)p * P.pos_stride] = tile[l * P.pitch + tix(p)];
    }
    if (LAST) {
        // block min / max (in the tile's first words, free after the store),
        // then one atomic per value, only when it improves the global one
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            omin = min(omin, (uint32_t)__shfl_xor((int)omin, d, 64));
            omax = max(omax, (uint32_t)__shfl_xor((int)omax, d, 64));
        }
        __syncthreads();
        if (lane == 0) {
            tile[2 * wv] = omin;
            tile[2 * wv + 1] = omax;
        }
        __syncthreads();
        if (tid == 0) {
            for (int w = 1; w < P.lines; ++w) {
                omin = min(omin, tile[2 * w]);
                omax = max(omax, tile[2 * w + 1]);
            }
            if (omin < __hip_atomic_load(P.minmax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMin(P.minmax, omin);
            if (omax > __hip_atomic_load(P.minmax + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMax(P.minmax + 1, omax);
        }
    }
}

__global__ void k_edt_minmax_init(uint32_t *mm) {
    if (threadIdx.x == 0) {
        mm[0] = 0xffffffffu;
        mm[1] = 0u;
    }
}

// Pass x: g1 = squared distance to the nearest site of the row (sites:
// data == 0, or data != 0 when INVERT), kNone when the row has none.  One
// wave per row: a ballot per 64-position chunk gives the chunk's site mask
// (lane c keeps chunk c's), prefix max / suffix min over the chunks give the
// nearest site before and after each chunk, and each position combines them
// with its own chunk's mask.  Reads 1 byte and writes 4 per voxel, both
// coalesced.
template <bool INVERT>
__global__ __launch_bounds__(256) void k_edt_rows(const uint8_t *__restrict__ data, int64_t rows, int nx,
                                                  uint32_t *__restrict__ out) {
    constexpr int kNoSite = 1 << 30;
    const int lane = threadIdx.x & 63;
    const int nch = (nx + 63) / 64;  // <= 64
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < rows; r += wstride) {
        const uint8_t *row = data + r * nx;
        uint32_t *orow = out + r * nx;
        uint32_t mlo = 0, mhi = 0;  // lane c: chunk c's site mask
        for (int c = 0; c < nch; ++c) {
            const int x = 64 * c + lane;
            bool site = false;
            if (x < nx) site = (row[x] == 0) != INVERT;
            const uint64_t m = __ballot(site);
            if (lane == c) {
                mlo = (uint32_t)m;
                mhi = (uint32_t)(m >> 32);
            }
        }
        const uint64_t mine = ((uint64_t)mhi << 32) | mlo;
        int last = lane < nch && mine != 0 ? 64 * lane + top_bit(mine) : -1;
        int first = lane < nch && mine != 0 ? 64 * lane + low_bit(mine) : kNoSite;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int a = __shfl_up(last, d, 64);
            const int b = __shfl_down(first, d, 64);
            if (lane >= d) last = max(last, a);
            if (lane + d < 64) first = min(first, b);
        }
        int before = __shfl_up(last, 1, 64);  // nearest site before chunk `lane`
        int after = __shfl_down(first, 1, 64);
        if (lane == 0) before = -1;
        if (lane == 63) after = kNoSite;
        for (int c = 0; c < nch; ++c) {
            const uint64_t m = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)mhi, c) << 32) |
                               (uint32_t)__builtin_amdgcn_readlane((int)mlo, c);
            const int bc = __builtin_amdgcn_readlane(before, c), ac = __builtin_amdgcn_readlane(after, c);
            const int x = 64 * c + lane;
            const uint64_t below = m & bits_below(lane + 1), above = m & ~bits_below(lane);
            const int l = below != 0 ? 64 * c + top_bit(below) : bc;
            const int nx_site = above != 0 ? 64 * c + low_bit(above) : ac;
            int d = kNoSite;
            if (l >= 0) d = x - l;
            if (nx_site < kNoSite) d = min(d, nx_site - x);
            if (x < nx) orow[x] = d < kNoSite ? (uint32_t)(d * d) : kNone;
        }
    }
}

// general.hh:96-118 on the float field: (uint16)((f - vmin) / (vmax - vmin) *
// 65535) with vmax starting at FLT_MIN (general.hh:78), then x cylinder.
// Rows of nx voxels, 4 per thread (one 16-byte load where nx % 4 == 0).
__global__ __launch_bounds__(256) void k_edt_to_u16(const float *__restrict__ f, int64_t rows, int64_t ny,
                                                    int64_t nx, const uint8_t *__restrict__ cyl,
                                                    const uint32_t *__restrict__ mm,
                                                    uint16_t *__restrict__ out) {
    const float vmin = ord2f_d(mm[0]);
    const float vmax = fmaxf(ord2f_d(mm[1]), FLT_MIN);
    const float span = vmax - vmin;
    const int64_t per_row = (nx + 1023) / 1024;
    const int64_t items = rows * per_row;
    const bool vec = nx % 4 == 0;
    for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
        const int64_t r = it / per_row;
        const int64_t x0 = (it % per_row) * 1024 + 4 * threadIdx.x;
        if (x0 >= nx) continue;
        const float *fr = f + r * nx;
        const uint8_t *cr = cyl != nullptr ? cyl + (r % ny) * nx : nullptr;
        uint16_t *orow = out + r * nx;
        float v[4];
        if (vec) {
            const float4 q = *reinterpret_cast<const float4 *>(fr + x0);
            v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k] = x0 + k < nx ? fr[x0 + k] : 0.0f;
        }
        uint16_t o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            o[k] = (uint16_t)((v[k] - vmin) / span * 65535.0f);
            if (cr != nullptr && x0 + k < nx && cr[x0 + k] == 0) o[k] = 0;
        }
        if (vec) {
            *reinterpret_cast<uint2 *>(orow + x0) =
                make_uint2((uint32_t)o[0] | ((uint32_t)o[1] << 16), (uint32_t)o[2] | ((uint32_t)o[3] << 16));
        } else {
            for (int k = 0; k < 4 && x0 + k < nx; ++k) orow[x0 + k] = o[k];
        }
    }
}

// Band length per lane: the least power of two with 64 L >= n.
int64_t band_of(int64_t n) {
    int64_t L = 1;
    while (64 * L < n) L <<= 1;
    return L;
}

size_t pass_smem(const EdtPass &P) {
    const int64_t L = band_of(P.n);
    return (size_t)4 * P.lines * ((size_t)P.pitch + (size_t)(64 * L) + kMetaInts);
}

template <int L, bool LAST>
hipError_t launch_l(const EdtPass &P, hipStream_t s) {
    if (P.lines > edt_max_lines(L)) return hipErrorInvalidValue;  // beyond the launch bound
    const size_t smem = pass_smem(P);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_edt_pass<L, LAST>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    const int64_t blocks = ((P.n_inner + P.lines - 1) / P.lines) * P.n_outer;
    hipLaunchKernelGGL((k_edt_pass<L, LAST>), dim3((unsigned)blocks), dim3(P.lines * 64), smem, s, P);
    return hipGetLastError();
}

template <bool LAST>
hipError_t launch_mode(const EdtPass &P, hipStream_t s) {
    const int64_t need = (P.n + 63) / 64;
    if (need <= 1) return launch_l<1, LAST>(P, s);
    if (need <= 2) return launch_l<2, LAST>(P, s);
    if (need <= 4) return launch_l<4, LAST>(P, s);
    if (need <= 8) return launch_l<8, LAST>(P, s);
    if (need <= 16) return launch_l<16, LAST>(P, s);
    if (need <= 32) return launch_l<32, LAST>(P, s);
    return launch_l<64, LAST>(P, s);
}

// Geometry of the pass along axis a (0 z, 1 y, 2 x) of an (nz, ny, nx) volume.
EdtPass pass_geom(int a, int64_t nz, int64_t ny, int64_t nx) {
    EdtPass P{};
    if (a == 2) {
        P.n = nx; P.pos_stride = 1;
        P.n_inner = nz * ny; P.inner_stride = nx; P.n_outer = 1; P.outer_stride = 0;
    } else if (a == 1) {
        P.n = ny; P.pos_stride = nx;
        P.n_inner = nx; P.inner_stride = 1; P.n_outer = nz; P.outer_stride = ny * nx;
    } else {
        P.n = nz; P.pos_stride = ny * nx;
        P.n_inner = ny * nx; P.inner_stride = 1; P.n_outer = 1; P.outer_stride = 0;
    }
    P.pitch = (int)(P.n + P.n / 16 + 1);
    // one wave per line; the largest count within the block budget
    int lines = edt_max_lines((int)band_of(P.n));
    while (lines > 1 && (int64_t)4 * lines * (P.pitch + 64 * band_of(P.n) + kMetaInts) > kBlockBudget)
        lines >>= 1;
    P.lines = lines;
    return P;
}

hipError_t run_pass(bool last, EdtPass P, hipStream_t s) {
    if (P.n_inner * P.n_outer == 0 || P.n == 0) return hipSuccess;
    return last ? launch_mode<true>(P, s) : launch_mode<false>(P, s);
}

hipError_t run_rows(const uint8_t *data, bool invert, int64_t rows, int64_t nx, uint32_t *out, hipStream_t s) {
    if (rows == 0 || nx == 0) return hipSuccess;
    int64_t blocks = (rows + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    if (invert)
        hipLaunchKernelGGL(k_edt_rows<true>, dim3((unsigned)blocks), dim3(256), 0, s, data, rows, (int)nx, out);
    else
        hipLaunchKernelGGL(k_edt_rows<false>, dim3((unsigned)blocks), dim3(256), 0, s, data, rows, (int)nx, out);
    return hipGetLastError();
}

int64_t vol_words_bytes(int64_t n) { return (n * 4 + 255) / 256 * 256; }

#define MXB_EHIP(expr)                                                          \
    do {                                                                        \
        hipError_t _e = (expr);                                                 \
        if (_e != hipSuccess) {                                                 \
            set_error(std::string(#expr) + ": " + hipGetErrorString(_e));       \
            return _e == hipErrorOutOfMemory ? ENOMEM_ : EDEVICE;               \
        }                                                                       \
    } while (0)

int check_extents(const char *who, int64_t nz, int64_t ny, int64_t nx) {
    if (nz < 0 || ny < 0 || nx < 0) {
        set_error(std::string(who) + ": negative extent");
        return EINVAL_;
    }
    if (nz > MXB_EDT_MAX_AXIS || ny > MXB_EDT_MAX_AXIS || nx > MXB_EDT_MAX_AXIS) {
        set_error(std::string(who) + ": at most " + std::to_string(MXB_EDT_MAX_AXIS) + " voxels per axis");
        return EINVAL_;
    }
    return OK;
}

// The three passes; the first reads `data` (sites: data == 0, or != 0 when
// `invert`), the last writes `last_out` in `last_mode` (2: squared, 3: float
// x cylinder + min/max into mm).
int device_edt_core(const uint8_t *data, int64_t nz, int64_t ny, int64_t nx, bool invert, uint32_t *a,
                    uint32_t *b, uint32_t *last_out, int last_mode, const uint8_t *cyl, uint32_t *mm,
                    hipStream_t s) {
    EdtPass py = pass_geom(1, nz, ny, nx), pz = pass_geom(0, nz, ny, nx);
    py.in = a; py.out = b;
    pz.in = b; pz.out = last_out; pz.cyl = cyl; pz.minmax = mm;
    if (last_mode == 3) {
        hipLaunchKernelGGL(k_edt_minmax_init, dim3(1), dim3(64), 0, s, mm);
        MXB_EHIP(hipGetLastError());
    }
    MXB_EHIP(run_rows(data, invert, nz * ny, nx, a, s));
    MXB_EHIP(run_pass(false, py, s));
    MXB_EHIP(run_pass(last_mode == 3, pz, s));
    return OK;
}

struct EBuf {
    void *p = nullptr;
    ~EBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t n) {
        hipError_t e = hipMalloc(&p, n > 0 ? n : 1);
        if (e != hipSuccess) {
            set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
            return e == hipErrorOutOfMemory ? ENOMEM_ : EDEVICE;
        }
        return OK;
    }
};

}  // namespace

}  // namespace mxb

using namespace mxb;

extern "C" {

int64_t mxb_edt_workspace_bytes(int64_t nz, int64_t ny, int64_t nx) {
    if (nz < 0 || ny < 0 || nx < 0) return -1;
    return 2 * vol_words_bytes(nz * ny * nx) + 256;
}

int mxb_device_edt(const uint8_t *dev_data, int64_t nz, int64_t ny, int64_t nx, uint32_t *dev_sq_out,
                   float *dev_out, void *dev_workspace, void *stream) {
    int rc = check_extents("edt", nz, ny, nx);
    if (rc != OK) return rc;
    const int64_t n = nz * ny * nx;
    if (n == 0) return OK;
    if (dev_data == nullptr || dev_workspace == nullptr || (dev_sq_out == nullptr && dev_out == nullptr)) {
        set_error("edt: null buffer");
        return EINVAL_;
    }
    hipStream_t s = (hipStream_t)stream;
    uint32_t *a = (uint32_t *)dev_workspace;
    uint32_t *b = (uint32_t *)((char *)dev_workspace + vol_words_bytes(n));
    uint32_t *mm = (uint32_t *)((char *)dev_workspace + 2 * vol_words_bytes(n));
    if (dev_sq_out != nullptr && (rc = device_edt_core(dev_data, nz, ny, nx, false, a, b, dev_sq_out, 2,
                                                       nullptr, mm, s)) != OK)
        return rc;
    if (dev_out != nullptr && (rc = device_edt_core(dev_data, nz, ny, nx, false, a, b, (uint32_t *)dev_out, 3,
                                                    nullptr, mm, s)) != OK)
        return rc;
    return OK;
}

int mxb_device_edt_field(const uint8_t *dev_implant_solid, int64_t nz, int64_t ny, int64_t nx,
                         const uint8_t *dev_cylinder, uint16_t *dev_out, void *dev_workspace, void *stream) {
    int rc = check_extents("edt_field", nz, ny, nx);
    if (rc != OK) return rc;
    const int64_t n = nz * ny * nx;
    if (n == 0) return OK;
    if (dev_implant_solid == nullptr || dev_out == nullptr || dev_workspace == nullptr) {
        set_error("edt_field: null buffer");
        return EINVAL_;
    }
    hipStream_t s = (hipStream_t)stream;
    uint32_t *a = (uint32_t *)dev_workspace;
    uint32_t *b = (uint32_t *)((char *)dev_workspace + vol_words_bytes(n));
    uint32_t *mm = (uint32_t *)((char *)dev_workspace + 2 * vol_words_bytes(n));
    // 0950:41 negates the implant mask: the sites are the implant voxels
    if ((rc = device_edt_core(dev_implant_solid, nz, ny, nx, true, a, b, a, 3, dev_cylinder, mm, s)) != OK)
        return rc;
    int64_t blocks = nz * ny * ((nx + 1023) / 1024);
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_edt_to_u16, dim3((unsigned)blocks), dim3(256), 0, s, (const float *)a, nz * ny, ny, nx,
                       dev_cylinder, mm, dev_out);
    MXB_EHIP(hipGetLastError());
    return OK;
}

int mxb_edt(const uint8_t *data, int64_t nz, int64_t ny, int64_t nx, uint32_t *sq_out, float *out,
            int verbose) {
    (void)verbose;
    int rc = check_extents("edt", nz, ny, nx);
    if (rc != OK) return rc;
    const int64_t n = nz * ny * nx;
    if (n == 0) return OK;
    if (data == nullptr || (sq_out == nullptr && out == nullptr)) {
        set_error("edt: null buffer");
        return EINVAL_;
    }
    EBuf d, ws, o2, of;
    if ((rc = d.alloc(n)) || (rc = ws.alloc(mxb_edt_workspace_bytes(nz, ny, nx)))) return rc;
    if (sq_out != nullptr && (rc = o2.alloc(n * 4))) return rc;
    if (out != nullptr && (rc = of.alloc(n * 4))) return rc;
    MXB_EHIP(staged_copy(d.p, data, n, true, nullptr));
    rc = mxb_device_edt((const uint8_t *)d.p, nz, ny, nx, (uint32_t *)o2.p, (float *)of.p, ws.p, nullptr);
    if (rc != OK) return rc;
    if (sq_out != nullptr) MXB_EHIP(staged_copy(sq_out, o2.p, n * 4, false, nullptr));
    if (out != nullptr) MXB_EHIP(staged_copy(out, of.p, n * 4, false, nullptr));
    return OK;
}

int mxb_edt_field(const uint8_t *implant_solid, int64_t nz, int64_t ny, int64_t nx, const uint8_t *cylinder,
                  uint16_t *out, int verbose) {
    (void)verbose;
    int rc = check_extents("edt_field", nz, ny, nx);
    if (rc != OK) return rc;
    const int64_t n = nz * ny * nx;
    if (n == 0) return OK;
    if (implant_solid == nullptr || out == nullptr) {
        set_error("edt_field: null buffer");
        return EINVAL_;
    }
    EBuf d, ws, o, c;
    if ((rc = d.alloc(n)) || (rc = ws.alloc(mxb_edt_workspace_bytes(nz, ny, nx))) || (rc = o.alloc(n * 2)))
        return rc;
    if (cylinder != nullptr && (rc = c.alloc(ny * nx))) return rc;
    MXB_EHIP(staged_copy(d.p, implant_solid, n, true, nullptr));
    if (cylinder != nullptr) MXB_EHIP(hipMemcpy(c.p, cylinder, ny * nx, hipMemcpyHostToDevice));
    rc = mxb_device_edt_field((const uint8_t *)d.p, nz, ny, nx, (const uint8_t *)c.p, (uint16_t *)o.p, ws.p,
                              nullptr);
    if (rc != OK) return rc;
    MXB_EHIP(staged_copy(out, o.p, n * 2, false, nullptr));
    return OK;
}

}  // extern "C"
