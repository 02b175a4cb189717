/* Ceiling probes of libmxb_diffusion.so: the achievable HBM and MFMA rates
 * the diffusion kernels are measured against.
 *
 * No counterpart in the reference (src/lib/cpp has no measurement kernels):
 * these entries exist so that the bandwidth and matrix-core ceilings quoted
 * next to the diffusion kernels' rooflines are reproducible from this
 * library.  They compute nothing the diffusion path uses.
 *
 * Both entries launch on `stream` (a hipStream_t, null = default stream),
 * time `reps` launches after one untimed launch with HIP events, wait for
 * them and write the mean milliseconds per launch.  They allocate nothing.
 * Return MXB_OK, MXB_EINVAL for bad arguments, MXB_EDEVICE for a HIP error
 * (codes of maxibone_diffusion.h).
 */
#ifndef MAXIBONE_PROBE_H
#define MAXIBONE_PROBE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum mxb_probe_hbm_kind {
    MXB_PROBE_READ = 0,    /* reads `bytes` of src (16-byte vectors)          */
    MXB_PROBE_WRITE = 1,   /* writes `bytes` of dst                           */
    MXB_PROBE_COPY = 2,    /* copies `bytes` from src to dst                  */
    MXB_PROBE_COPY_NT = 3  /* the same copy with non-temporal loads and stores */
};

/* `bytes` a positive multiple of 16384; src and dst 16-byte aligned device
 * buffers of at least `bytes` (src unused by MXB_PROBE_WRITE).  MXB_PROBE_READ
 * writes dst's first 16 bytes only if the xor of some thread's 16 words is
 * 0x9e3779b9, which a zero-filled src never gives. */
int mxb_probe_hbm(int kind, const void *src, void *dst, int64_t bytes, int reps, void *stream,
                  float *ms_per_launch);

enum mxb_probe_mfma_kind {
    MXB_PROBE_MFMA_F32 = 0,  /* v_mfma_f32_16x16x4f32      */
    MXB_PROBE_MFMA_BF16 = 1  /* v_mfma_f32_16x16x32_bf16   */
};

/* 4096 waves each issue `iters` x 8 independent MFMAs; `sink` is a device
 * buffer of at least 4 bytes (left unchanged).  *flops_per_launch receives
 * the dense flops one launch performs (2 M N K per MFMA). */
int mxb_probe_mfma(int kind, int64_t iters, void *sink, int reps, void *stream, float *ms_per_launch,
                   double *flops_per_launch);

#ifdef __cplusplus
}
#endif

#endif
