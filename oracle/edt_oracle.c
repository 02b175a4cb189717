/*
 * Exact squared Euclidean distance transform of a binary volume: the oracle
 * for the EDT field of src/processing_steps/0950_generate_edt.py.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline); the product path never calls it.
 *
 * 0950:49 calls edt.edt(mask, parallel=n) of the external `edt` package
 * (seung-lab edt, pinned at 2.4.0 in src/requirements.txt:2; not present in
 * the reference tree).  On a boolean volume with its defaults (anisotropy
 * (1, 1, 1), black_border=False) that is the Euclidean distance transform:
 * every nonzero voxel gets the distance to the nearest zero voxel, zero
 * voxels get 0, and the volume edge is not a boundary.  The package computes
 * squared distances separably (a 1D pass along the first axis, then the
 * lower envelope of parabolas of Felzenszwalb & Huttenlocher, "Distance
 * Transforms of Sampled Functions", Theory of Computing 8 (2012) 415-428,
 * along the others) and returns the float32 square root.
 *
 * This restatement does the same in exact integer arithmetic:
 *   pass x:    g1[v] = (x - x*)^2 for the nearest zero x* of the row
 *              (two scans), "none" = OEDT_INF;
 *   pass y, z: g'[i] = min_j g[j] + (i - j)^2 by Algorithm 1 of the paper,
 *              with the parabola intersections compared as exact fractions
 *              (128-bit cross products) instead of floating point;
 *   result:    d2 (uint64, OEDT_INF where the volume has no zero voxel).
 * Its squared distances are therefore the mathematical ones; the tests pin
 * it against brute force and scipy.ndimage.distance_transform_edt.
 */
#include <stdint.h>
#include <stdlib.h>

#define OEDT_INF UINT64_MAX

/* 1D lower envelope of parabolas g[j] + (i - j)^2 over one line of n samples
 * (stride `st`), written to out (same stride).  v, scratch of n entries. */
static void fh_line(const uint64_t *g, uint64_t *out, int64_t n, int64_t st, int64_t *v) {
    int64_t k = -1;
    for (int64_t q = 0; q < n; ++q) {
        const uint64_t gq = g[q * st];
        if (gq == OEDT_INF) continue;
        const __int128 cq = (__int128)gq + (__int128)q * q;
        /* pop v[k] while the new parabola's intersection with v[k] lies at or
         * left of v[k]'s intersection with v[k-1]:
         *   s(v[k], q) <= s(v[k-1], v[k]), s(a, b) = (c_b - c_a) / (2 (b - a)) */
        while (k >= 1) {
            const int64_t a = v[k - 1], b = v[k];
            const __int128 ca = (__int128)g[a * st] + (__int128)a * a;
            const __int128 cb = (__int128)g[b * st] + (__int128)b * b;
            if ((cq - cb) * (b - a) <= (cb - ca) * (q - b))
                --k;
            else
                break;
        }
        v[++k] = q;
    }
    if (k < 0) {
        for (int64_t i = 0; i < n; ++i) out[i * st] = OEDT_INF;
        return;
    }
    int64_t j = 0;
    for (int64_t i = 0; i < n; ++i) {
        /* advance while the boundary s(v[j], v[j+1]) < i */
        while (j < k) {
            const int64_t a = v[j], b = v[j + 1];
            const __int128 ca = (__int128)g[a * st] + (__int128)a * a;
            const __int128 cb = (__int128)g[b * st] + (__int128)b * b;
            if (cb - ca < (__int128)2 * i * (b - a))
                ++j;
            else
                break;
        }
        const int64_t a = v[j];
        out[i * st] = g[a * st] + (uint64_t)((i - a) * (i - a));
    }
}

/* d2[z][y][x]: squared distance of voxel (z, y, x) to the nearest zero of
 * m (C order).  Returns 0, or -1 when out of memory. */
int oracle_edt_sq(const uint8_t *m, int64_t nz, int64_t ny, int64_t nx, uint64_t *d2) {
    const int64_t n = nz * ny * nx;
    if (n == 0) return 0;
    int64_t nmax = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;
    uint64_t *tmp = malloc((size_t)n * sizeof(uint64_t));
    int64_t *v = malloc((size_t)nmax * sizeof(int64_t));
    if (tmp == NULL || v == NULL) {
        free(tmp);
        free(v);
        return -1;
    }
    /* pass x: nearest zero of the row, by a forward and a backward scan */
    for (int64_t r = 0; r < nz * ny; ++r) {
        const uint8_t *row = m + r * nx;
        uint64_t *o = d2 + r * nx;
        int64_t last = -1;
        for (int64_t x = 0; x < nx; ++x) {
            if (row[x] == 0) last = x;
            o[x] = last < 0 ? OEDT_INF : (uint64_t)((x - last) * (x - last));
        }
        last = -1;
        for (int64_t x = nx - 1; x >= 0; --x) {
            if (row[x] == 0) last = x;
            if (last >= 0) {
                const uint64_t d = (uint64_t)((last - x) * (last - x));
                if (d < o[x]) o[x] = d;
            }
        }
    }
    /* pass y: d2 -> tmp */
    for (int64_t z = 0; z < nz; ++z)
        for (int64_t x = 0; x < nx; ++x)
            fh_line(d2 + z * ny * nx + x, tmp + z * ny * nx + x, ny, nx, v);
    /* pass z: tmp -> d2 */
    for (int64_t y = 0; y < ny; ++y)
        for (int64_t x = 0; x < nx; ++x)
            fh_line(tmp + y * nx + x, d2 + y * nx + x, nz, ny * nx, v);
    free(tmp);
    free(v);
    return 0;
}
