/*
 * maxibone_diffusion.h - C ABI of the MI355X (gfx950) masked 3D diffusion.
 *
 * The diffusion (reference: src/lib/cpp/cpu/diffusion.cc, semantics of
 * cpu:339-374): u0 = (mask > 0); per repetition correlate along z, then y,
 * then x with the 1D kernel (zero outside the volume), then set mask voxels
 * back to 1; finally floor(u * 65535) as uint16.
 *
 * Volumes are C-order (z, y, x), x fastest.  The kernel has K = 2R + 1 taps
 * (K odd).  Every function returns an mxb_status; on failure
 * mxb_last_error() describes the cause (thread-local, valid until the next
 * call on the same thread).
 *
 * Pointers named host_* / plain `mask`, `kernel`, `output` are host memory;
 * pointers named dev_* are device (HBM) memory of the current HIP device.
 */
#ifndef MAXIBONE_DIFFUSION_H
#define MAXIBONE_DIFFUSION_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MXB_OK = 0,
    MXB_EINVAL = 1,  /* invalid argument (Python: ValueError)          */
    MXB_EDEVICE = 2, /* HIP runtime / kernel failure (RuntimeError)    */
    MXB_ENOMEM = 3,  /* device or host allocation failed (MemoryError) */
    MXB_EIO = 4      /* file I/O failed (RuntimeError)                 */
} mxb_status;

/* Storage / arithmetic form of the diffusion state. */
typedef enum {
    MXB_FP32 = 0,       /* FP32 state, FP32 banded-Toeplitz MFMA passes          */
    MXB_BF16 = 1,       /* BF16 state, BF16 MFMA passes, FP32-accuracy final
                           repetition (see below)                               */
    MXB_FP16 = 2,       /* FP16 state, FP16 MFMA passes, FP32-accuracy final
                           repetition: in the MXB_FP32_SPLIT form where its
                           fused x -> y pass serves the volume (x extent a
                           multiple of 8), else in MXB_FP32; a single
                           repetition runs in MXB_FP32                          */
    MXB_FP32_FMA = 3,   /* FP32 state, one VALU FMA per tap                       */
    MXB_FP32_EXACT = 4, /* FP32 state, product-then-add: bit-equal to the
                           reference CPU build                                  */
    MXB_FP32_SPLIT = 5  /* FP32-class state held as FP16 hi + lo (u = hi + lo,
                           22 significant bits); each pass runs three FP16
                           MFMAs per chunk (hi*k_hi + hi*k_lo + lo*k_hi) on
                           power-of-two-scaled taps with FP32 accumulation.
                           Same HBM bytes as FP32, a fraction of its MFMA
                           time; within the FP32 bar (1e-5, 1 uint16 LSB) of
                           the reference CPU path.  Volumes whose x extent
                           is not a multiple of 8, and radii above 64, run
                           MXB_FP32.                                       */
} mxb_precision;

/* ---- drop-in entry points (host buffers) ------------------------------- */

/* Replaces python_api::diffusion_in_memory (src/pybind/diffusion-pybind.cc:27-53)
 * and NS::diffusion_in_memory (src/lib/cpp/include/diffusion.hh:36,
 * cpu:339-374, gpu:758-806): uint8 mask in, uint16 result out, FP32 state.
 * When the volume does not fit in device memory this streams z-chunks through
 * the device, like the reference's automatic switch to
 * diffusion_out_of_core (pyb:40-45).  The environment variable
 * MXB_DEVICE_BUDGET (bytes) lowers the device memory the switch assumes. */
int mxb_diffusion_u8_u16(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx,
                         const float *kernel, int64_t kernel_size, int64_t repetitions,
                         uint16_t *output, int verbose);

/* Same as mxb_diffusion_u8_u16 but returns the FP32 state before
 * quantisation (no reference counterpart: the reference only exposes the
 * uint16 result, cpu:365). */
int mxb_diffusion_u8_f32(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx,
                         const float *kernel, int64_t kernel_size, int64_t repetitions,
                         float *output, int verbose);

/* Extended form of the two above: any mxb_precision; either output may be
 * NULL (not both).  precision = MXB_FP32 is what mxb_diffusion_u8_u16 runs. */
int mxb_diffusion_ex(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx,
                     const float *kernel, int64_t kernel_size, int64_t repetitions,
                     int precision, uint16_t *output_u16, float *output_f32, int verbose);

/* z-chunked form over host-resident state: the device holds one chunk of
 * chunk_z layers plus a halo; `temporal` repetitions run per chunk visit with
 * a halo of temporal*R layers (temporal = 1 is the reference's
 * one-step-per-chunk scheme, cpu:433-453 / gpu:869-905).  Results equal the
 * in-memory form of the same precision bit for bit: BF16 / FP16 keep 16-bit
 * host state between chunk visits and run the last repetition at FP32
 * accuracy, in the same form as the in-memory run.
 * Replaces NS::diffusion_out_of_core (hh:74, gpu:947-1129). */
int mxb_diffusion_chunked(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx,
                          const float *kernel, int64_t kernel_size, int64_t repetitions,
                          int64_t chunk_z, int64_t temporal, int precision,
                          uint16_t *output_u16, float *output_f32, int verbose);

/* File-backed form.  Replaces python_api::diffusion_on_disk (pyb:67-77, bound
 * at pyb:85) and NS::diffusion_on_disk (hh:55, cpu:376-468, gpu:808-945):
 * input_file holds raw uint8 (z, y, x) of total_shape; output_file receives
 * raw uint16 (the GPU build's output type, gpu:937); global_shape is the
 * chunk shape (its y and x must equal total_shape's). */
int mxb_diffusion_on_disk(const char *input_file, const float *kernel, int64_t kernel_size,
                          const char *output_file, int64_t total_z, int64_t total_y,
                          int64_t total_x, int64_t global_z, int64_t global_y, int64_t global_x,
                          int64_t repetitions, int verbose);

/* Flags of mxb_diffusion_on_disk_ex. */
enum {
    /* Illuminate, at every chunk visit, the voxels whose incoming state is
     * exactly 1.0f instead of the input mask: the reference's on-disk rule
     * (store_mask, src/lib/cpp/cpu/diffusion.cc:331-337 called at :445;
     * src/lib/cpp/gpu/diffusion.cc:745-756 called at :885).  Runs one
     * repetition per chunk visit, as the reference does, and needs an FP32
     * precision.  Without it the result equals the in-memory form. */
    MXB_ON_DISK_REFERENCE_MASK = 1
};

/* mxb_diffusion_on_disk with a precision (mxb_precision; the 16-bit forms keep
 * 16-bit state files and run their last repetition at FP32 accuracy) and flags (0 or
 * MXB_ON_DISK_REFERENCE_MASK).  Replaces the same reference interface as
 * mxb_diffusion_on_disk (pyb:67-77; cpu:376-468; gpu:808-945). */
int mxb_diffusion_on_disk_ex(const char *input_file, const float *kernel, int64_t kernel_size,
                             const char *output_file, int64_t total_z, int64_t total_y,
                             int64_t total_x, int64_t global_z, int64_t global_y, int64_t global_x,
                             int64_t repetitions, int precision, int flags, int verbose);

/* ---- batched form --------------------------------------------------------- */

/* One independent sub-volume of a batch (no reference counterpart: the
 * reference diffuses one volume per call). */
typedef struct {
    const uint8_t *mask;   /* host, (nz, ny, nx) C-order   */
    int64_t nz, ny, nx;
    const float *kernel;   /* host, kernel_size taps        */
    int64_t kernel_size;
    uint16_t *output_u16;  /* host, may be NULL              */
    float *output_f32;     /* host, may be NULL              */
} mxb_volume_desc;

/* Diffuse n volumes, each with its own kernel, `repetitions` times, in one
 * persistent launch per pass, with the repetition loop captured in a HIP
 * graph.  (No reference counterpart.) */
int mxb_diffusion_batched(const mxb_volume_desc *volumes, int n, int64_t repetitions,
                          int precision, int verbose);

/* Device-resident batch.  A plan pads each volume into a pitched workspace
 * (rows of 128 bytes of state), builds the per-pass work lists, autotunes the
 * persistent kernel's grid and segment length (the fastest of three timed
 * repetitions per candidate; cached per batch signature, and in the file
 * named by MXB_TUNE_FILE), and captures the whole repetition loop in a HIP
 * graph.  (No reference counterpart.) */
typedef struct {
    const uint8_t *dev_mask;  /* device, (nz, ny, nx) C-order */
    int64_t nz, ny, nx;
    const float *kernel;      /* host, kernel_size taps       */
    int64_t kernel_size;
    uint16_t *dev_out_u16;    /* device, may be NULL           */
    float *dev_out_f32;       /* device, may be NULL           */
} mxb_device_volume;

typedef struct mxb_batch_plan mxb_batch_plan;

/* stream: a hipStream_t; NULL makes the plan create its own stream. */
int mxb_batch_create(const mxb_device_volume *volumes, int n, int64_t repetitions, int precision,
                     void *stream, mxb_batch_plan **plan);
/* mode 0: the captured graph of persistent passes (FP32, and BF16 / FP16
 * with their FP32 last repetition; bit-identical to mode 1), when every
 * radius is at most 64 (else, and for the FMA / exact FP32 forms, mode 1);
 * mode 1: per-volume launches of the single-volume path (any precision). */
int mxb_batch_run(mxb_batch_plan *plan, int mode);
/* Wait for the plan's stream. */
int mxb_batch_sync(mxb_batch_plan *plan);
int mxb_batch_info(const mxb_batch_plan *plan, int *persistent, int *grid, int *seg,
                   double *tuned_ms);
int mxb_batch_destroy(mxb_batch_plan *plan);

/* ---- device-resident form (inputs already in HBM) -------------------------- */

/* Bytes of one state buffer for `precision` (two are needed): 4 bytes for
 * each of the voxel count rounded up to 128 plus 2048 (the split form's low
 * plane starts 4 KB past its high plane's end), then room for the mask packed
 * to bits (one 32-bit word per 32 voxels of an x row), which the first buffer
 * carries.  -1 (and mxb_last_error set) for an unknown precision or a
 * negative extent. */
int64_t mxb_state_bytes(int precision, int64_t nz, int64_t ny, int64_t nx);

/* Run the diffusion on device buffers, enqueued on `stream` (a hipStream_t;
 * NULL = default stream).  dev_state0/1: two buffers of mxb_state_bytes each.
 * dev_out_u16 / dev_out_f32: final outputs (either may be NULL, not both).
 * Allocates nothing.  Returns after enqueueing (asynchronous). */
int mxb_device_diffusion(const uint8_t *dev_mask, int64_t nz, int64_t ny, int64_t nx,
                         const float *kernel, int64_t kernel_size, int64_t repetitions,
                         int precision, void *dev_state0, void *dev_state1,
                         uint16_t *dev_out_u16, float *dev_out_f32, void *stream);

/* Bytes of device memory that mxb_device_diffusion_noalloc holds for its state
 * buffers on every device it runs on (20 MiB). */
#define MXB_NOALLOC_SCRATCH_BYTES 20971520

/* mxb_device_diffusion without the state buffers: no reference counterpart.
 * It is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The two
 * state buffers live in a fixed device array of MXB_NOALLOC_SCRATCH_BYTES per
 * device, held until the process ends, so the volume's two state buffers
 * (2 * mxb_state_bytes, the second rounded up to 256 bytes) must fit in it:
 * every volume of at most 128^3 voxels with an x extent of at least 32 does.
 * Larger volumes fail with MXB_EINVAL before any device work.  Allocates no
 * device or host memory and records no timing events.  Every call on a device
 * uses that device's one array: calls that overlap on one device (other
 * threads, other streams) race; calls ordered on one stream do not.
 * Asynchronous, like mxb_device_diffusion. */
int mxb_device_diffusion_noalloc(const uint8_t *dev_mask, int64_t nz, int64_t ny, int64_t nx,
                                 const float *kernel, int64_t kernel_size, int64_t repetitions,
                                 int precision, uint16_t *dev_out_u16, float *dev_out_f32,
                                 void *stream);

/* ---- per-kernel timing (HIP events on the launch stream) ------------------ */

/* Kernel classes reported by mxb_kernel_times (MXB_K_XYPASS: the fused
 * x -> y pass of the 16-bit and split forms). */
typedef enum {
    MXB_K_ZPASS = 0,
    MXB_K_YPASS = 1,
    MXB_K_XPASS = 2,
    MXB_K_OTHER = 3,
    MXB_K_XYPASS = 4
} mxb_kernel_class;

/* Enable (1) / disable (0) event timing of every pass launch; resets totals. */
int mxb_set_kernel_timing(int enable);
/* Total milliseconds and launch count per class since timing was enabled
 * (synchronises the recorded events first). */
int mxb_kernel_times(int kernel_class, double *total_ms, int64_t *launches);

/* ---- misc --------------------------------------------------------------- */

/* Number of visible HIP devices (0 when none or the runtime is unusable). */
int mxb_device_count(void);
const char *mxb_last_error(void);
const char *mxb_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MAXIBONE_DIFFUSION_H */
