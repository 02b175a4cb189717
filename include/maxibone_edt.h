/* The implant EDT field on MI355X: the Euclidean distance transform that
 * src/processing_steps/0950_generate_edt.py computes with the external `edt`
 * package (seung-lab edt 2.4.0, src/requirements.txt:2), and 0950's
 * cylinder-mask and uint16 conversion around it.
 *
 * Every function returns an int status (MXB_OK = 0, else an mxb_status from
 * maxibone_diffusion.h; the message is in mxb_last_error()).  Host-buffer
 * entries copy in, compute on the GPU and copy out; device entries take HBM
 * pointers of the library's HIP runtime and a caller stream.
 *
 * Semantics (edt.edt(data) with its defaults: anisotropy (1, 1, 1),
 * black_border=False): every nonzero voxel of `data` gets the Euclidean
 * distance to the nearest zero voxel, zero voxels get 0, and the volume edge
 * is not a boundary.  Squared distances are exact integers; the float form
 * is the float32 square root of the float32 squared distance.  A volume with
 * no zero voxel gets MXB_EDT_NONE (squared) / +inf (float).  Binary input only
 * (0950 passes a boolean mask; the package's multi-label mode is not
 * provided).  Each axis may hold at most MXB_EDT_MAX_AXIS voxels.
 */
#ifndef MAXIBONE_EDT_H
#define MAXIBONE_EDT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MXB_EDT_NONE 0xffffffffu
#define MXB_EDT_MAX_AXIS 4096

/* edt.edt(data) (0950:49): squared distances (uint32) and/or distances
 * (float32); either output may be NULL, not both. */
int mxb_edt(const uint8_t *data, int64_t nz, int64_t ny, int64_t nx, uint32_t *sq_out,
            float *out, int verbose);

/* 0950:41-66 as one call: implant_mask = ~implant_solid; fedt = edt(implant_mask);
 * fedt *= cyl; fedt = to_int(fedt, uint16) (general.hh:96-118, the C++
 * normalized_convert, with its maximum starting at FLT_MIN, general.hh:78);
 * fedt *= cyl.  `implant_solid`: uint8 (nonzero = implant); `cylinder`: a
 * (ny, nx) uint8 mask (nonzero = inside), NULL = all inside (0950:51 uses
 * generate_cylinder_mask(nx)). */
int mxb_edt_field(const uint8_t *implant_solid, int64_t nz, int64_t ny, int64_t nx,
                  const uint8_t *cylinder, uint16_t *out, int verbose);

/* Device forms.  The workspace holds two uint32 volumes and the reduction
 * words; the calls are enqueued on `stream` and return without waiting. */
int64_t mxb_edt_workspace_bytes(int64_t nz, int64_t ny, int64_t nx);
int mxb_device_edt(const uint8_t *dev_data, int64_t nz, int64_t ny, int64_t nx,
                   uint32_t *dev_sq_out, float *dev_out, void *dev_workspace, void *stream);
int mxb_device_edt_field(const uint8_t *dev_implant_solid, int64_t nz, int64_t ny, int64_t nx,
                         const uint8_t *dev_cylinder, uint16_t *dev_out, void *dev_workspace,
                         void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MAXIBONE_EDT_H */
