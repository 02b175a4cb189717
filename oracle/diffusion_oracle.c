/*
 * CPU restatement of maxibone's masked 3D diffusion: the oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline); the product path never calls it.
 *
 * It restates, in plain C with OpenMP, the reference CPU path
 * (src/lib/cpp/cpu/diffusion.cc):
 *
 *   u0[v]  = M[v] > 0 ? 1 : 0                                  cpu:243-248
 *   per repetition (cpu:353-359, diffusion_step cpu:314-320):
 *     a = C_z(u); b = C_y(a); c = C_x(b)                       cpu:315-318
 *     C_d(f)[i] = sum_{r=max(-R,-i_d)}^{min(R,N_d-1-i_d)} f[i+r e_d]*k[R+r]
 *                 (sum starts at 0, r ascending, product rounded
 *                  to the storage type, then added: no FMA)   cpu:31-60
 *     u[v] = M[v] > 0 ? 1 : c[v]                               cpu:295-302
 *   out[v] = (uint16) floor(u[v] * 65535.0f)                   cpu:191-196
 *
 * Build with -ffp-contract=off so that no FMA is formed; the reference is
 * built with -O3 and no -march (src/Makefile:9-11), i.e. x86-64 baseline SSE2
 * without FMA, so the float variant reproduces it bit for bit.
 *
 * The double variant runs the same algorithm in FP64 (the FP64 oracle for the
 * reduced-precision forms); its taps are supplied by the caller as double.
 *
 * The chunked variant follows the z-chunk loop of diffusion_on_disk
 * (cpu:376-468: chunk of gz layers plus a halo of R layers per side, zero at
 * the volume ends, one full step per chunk per repetition, interior written
 * back) but illuminates with the true input mask instead of re-deriving the
 * mask as (u == 1.0f) (cpu:331-337, cpu:445); with that choice it equals the
 * in-memory result exactly.  `quirk_mask` = 1 restores the reference's
 * (u == 1.0f) rule.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MIN(a, b) ((a) < (b) ? (a) : (b))

/* cpu:243-248 */
void oracle_u8_to_f32(const uint8_t *m, float *u, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) u[i] = m[i] > 0 ? 1.0f : 0.0f;
}

/* cpu:191-196: multiply in float, floor, truncate to uint16. */
void oracle_quantise_u16(const float *u, uint16_t *q, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) q[i] = (uint16_t)floorf(u[i] * 65535.0f);
}

/* cpu:139-144 (used by the reference's on-disk output). */
void oracle_quantise_u8(const float *u, uint8_t *q, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) q[i] = (uint8_t)floorf(u[i] * 255.0f);
}

/* cpu:31-60.  dim 0 = z (stride ny*nx), 1 = y (stride nx), 2 = x (stride 1). */
void oracle_core_f32(const float *in, const float *k, float *out, int64_t nz,
                     int64_t ny, int64_t nx, int dim, int64_t R) {
    const int64_t stride[3] = {ny * nx, nx, 1};
    const int64_t Ns[3] = {nz, ny, nx};
#pragma omp parallel for collapse(3) schedule(static)
    for (int64_t i = 0; i < nz; i++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t l = 0; l < nx; l++) {
                const int64_t X[3] = {i, j, l};
                const int64_t lo = -MIN(R, X[dim]);
                const int64_t hi = MIN(R, Ns[dim] - X[dim] - 1);
                const int64_t o = i * stride[0] + j * stride[1] + l;
                float sum = 0.0f;
                for (int64_t r = lo; r <= hi; r++) {
                    const float p = in[o + r * stride[dim]] * k[R + r];
                    sum = sum + p;
                }
                out[o] = sum;
            }
}

void oracle_core_f64(const double *in, const double *k, double *out, int64_t nz,
                     int64_t ny, int64_t nx, int dim, int64_t R) {
    const int64_t stride[3] = {ny * nx, nx, 1};
    const int64_t Ns[3] = {nz, ny, nx};
#pragma omp parallel for collapse(3) schedule(static)
    for (int64_t i = 0; i < nz; i++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t l = 0; l < nx; l++) {
                const int64_t X[3] = {i, j, l};
                const int64_t lo = -MIN(R, X[dim]);
                const int64_t hi = MIN(R, Ns[dim] - X[dim] - 1);
                const int64_t o = i * stride[0] + j * stride[1] + l;
                double sum = 0.0;
                for (int64_t r = lo; r <= hi; r++) sum = sum + in[o + r * stride[dim]] * k[R + r];
                out[o] = sum;
            }
}

/* cpu:295-302 */
static void illuminate_f32(const uint8_t *m, float *u, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        if (m[i] > 0) u[i] = 1.0f;
}

static void illuminate_f64(const uint8_t *m, double *u, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        if (m[i] > 0) u[i] = 1.0;
}

/* cpu:314-320: z, y, x passes with ping-pong, then illuminate.  On return the
 * state is in b[0]. */
static void step_f32(const uint8_t *m, float **b, int64_t nz, int64_t ny, int64_t nx,
                     const float *k, int64_t R) {
    for (int dim = 0; dim < 3; dim++) {
        oracle_core_f32(b[0], k, b[1], nz, ny, nx, dim, R);
        float *t = b[0];
        b[0] = b[1];
        b[1] = t;
    }
    illuminate_f32(m, b[0], nz * ny * nx);
}

static void step_f64(const uint8_t *m, double **b, int64_t nz, int64_t ny, int64_t nx,
                     const double *k, int64_t R) {
    for (int dim = 0; dim < 3; dim++) {
        oracle_core_f64(b[0], k, b[1], nz, ny, nx, dim, R);
        double *t = b[0];
        b[0] = b[1];
        b[1] = t;
    }
    illuminate_f64(m, b[0], nz * ny * nx);
}

/* cpu:339-374.  Writes the float state to u_out (may be NULL) and its uint16
 * quantisation to q_out (may be NULL).  Returns 0, or -1 on bad arguments /
 * allocation failure. */
int oracle_diffusion_f32(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx,
                         const float *k, int64_t K, int64_t reps, float *u_out,
                         uint16_t *q_out) {
    if (nz < 0 || ny < 0 || nx < 0 || K < 1 || (K % 2) == 0 || reps < 0) return -1;
    const int64_t n = nz * ny * nx, R = K / 2;
    float *b[2];
    b[0] = (float *)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    b[1] = (float *)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    if (!b[0] || !b[1]) {
        free(b[0]);
        free(b[1]);
        return -1;
    }
    oracle_u8_to_f32(mask, b[0], n);
    for (int64_t t = 0; t < reps; t++) step_f32(mask, b, nz, ny, nx, k, R);
    if (u_out) memcpy(u_out, b[0], (size_t)n * sizeof(float));
    if (q_out) oracle_quantise_u16(b[0], q_out, n);
    free(b[0]);
    free(b[1]);
    return 0;
}

int oracle_diffusion_f64(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx,
                         const double *k, int64_t K, int64_t reps, double *u_out) {
    if (nz < 0 || ny < 0 || nx < 0 || K < 1 || (K % 2) == 0 || reps < 0) return -1;
    const int64_t n = nz * ny * nx, R = K / 2;
    double *b[2];
    b[0] = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    b[1] = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    if (!b[0] || !b[1]) {
        free(b[0]);
        free(b[1]);
        return -1;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) b[0][i] = mask[i] > 0 ? 1.0 : 0.0;
    for (int64_t t = 0; t < reps; t++) step_f64(mask, b, nz, ny, nx, k, R);
    memcpy(u_out, b[0], (size_t)n * sizeof(double));
    free(b[0]);
    free(b[1]);
    return 0;
}

/* Chunked z-slab restatement of cpu:376-468 on host memory instead of files.
 * gz = layers per chunk; each chunk is loaded with R halo layers on both
 * sides (load_partial, cpu:64-91: zero outside the volume), stepped once,
 * and its interior stored (store_partial, cpu:93-97) into the other state
 * buffer; the buffers swap after each repetition (cpu:451-452). */
int oracle_diffusion_chunked_f32(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx,
                                 const float *k, int64_t K, int64_t reps, int64_t gz,
                                 int quirk_mask, float *u_out) {
    if (nz < 0 || ny < 0 || nx < 0 || K < 1 || (K % 2) == 0 || reps < 0 || gz < 1) return -1;
    const int64_t n = nz * ny * nx, R = K / 2, layer = ny * nx;
    const int64_t pz = gz + 2 * R, pn = pz * layer;
    float *state[2], *cb[2];
    uint8_t *cm;
    state[0] = (float *)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    state[1] = (float *)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    cb[0] = (float *)malloc((size_t)(pn > 0 ? pn : 1) * sizeof(float));
    cb[1] = (float *)malloc((size_t)(pn > 0 ? pn : 1) * sizeof(float));
    cm = (uint8_t *)malloc((size_t)(pn > 0 ? pn : 1));
    if (!state[0] || !state[1] || !cb[0] || !cb[1] || !cm) {
        free(state[0]); free(state[1]); free(cb[0]); free(cb[1]); free(cm);
        return -1;
    }
    oracle_u8_to_f32(mask, state[0], n);
    const int64_t nblocks = (nz + gz - 1) / gz;
    for (int64_t t = 0; t < reps; t++) {
        for (int64_t blk = 0; blk < nblocks; blk++) {
            const int64_t z0 = blk * gz, zsz = MIN(gz, nz - z0);
            /* chunk buffer covers global layers [z0 - R, z0 - R + pz) */
            for (int64_t c = 0; c < pz; c++) {
                const int64_t zg = z0 - R + c;
                float *dst = cb[0] + c * layer;
                uint8_t *mdst = cm + c * layer;
                if (zg < 0 || zg >= nz || c >= zsz + 2 * R) {
                    memset(dst, 0, (size_t)layer * sizeof(float));
                    memset(mdst, 0, (size_t)layer);
                } else {
                    memcpy(dst, state[0] + zg * layer, (size_t)layer * sizeof(float));
                    if (quirk_mask) {
                        for (int64_t i = 0; i < layer; i++) mdst[i] = dst[i] == 1.0f;
                    } else {
                        memcpy(mdst, mask + zg * layer, (size_t)layer);
                    }
                }
            }
            float *bb[2] = {cb[0], cb[1]};
            step_f32(cm, bb, pz, ny, nx, k, R);
            memcpy(state[1] + z0 * layer, bb[0] + R * layer, (size_t)(zsz * layer) * sizeof(float));
        }
        float *s = state[0];
        state[0] = state[1];
        state[1] = s;
    }
    memcpy(u_out, state[0], (size_t)n * sizeof(float));
    free(state[0]); free(state[1]); free(cb[0]); free(cb[1]); free(cm);
    return 0;
}
