/* Consumers of the diffusion field on MI355X: the implant field combination
 * and the 2D histograms the analysis computes from the uint16 fields.
 *
 * Every function returns an int status (MXB_OK = 0, else an mxb_status from
 * maxibone_diffusion.h; the message is in mxb_last_error()).  Host-buffer
 * entries copy in, compute on the GPU and copy out; device entries take HBM
 * pointers of the library's HIP runtime and a caller stream.
 *
 * Semantics follow the reference CPU forms:
 *   combine:    src/processing_steps/0975_combine_fields.py:41-62 with the
 *               C++ to_int (src/lib/cpp/include/general.hh:96-118)
 *   histograms: src/lib/cpp/cpu/histograms.cc:14-78 (axis), :231-306 (field);
 *               bindings src/pybind/histograms-pybind.cc:16-108
 * with two documented differences: bins the reference would index out of
 * range are skipped, and the voxel bin is floored as in the CPU forms (the
 * reference's OpenACC axis histogram rounds it, src/lib/cpp/gpu/histograms.cc:94).
 */
#ifndef MAXIBONE_FIELDS_H
#define MAXIBONE_FIELDS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 0975:41-62: out = cyl * to_int_u16(cyl * rescale(gauss/max(gauss) -
 * edt/max(edt))), all arithmetic in float32 in the script's order.
 * `cylinder` is a (ny, nx) uint8 mask (nonzero = inside), NULL = all inside
 * (0975:39 uses generate_cylinder_mask(nx)).  Both fields must have a nonzero
 * maximum.  Replaces the numpy block of 0975:41-62 (no reference C entry). */
int mxb_combine_fields(const uint16_t *gauss, const uint16_t *edt, int64_t nz, int64_t ny,
                       int64_t nx, const uint8_t *cylinder, uint16_t *out, int verbose);

/* Device form: bytes of the workspace it needs (a few KB).  The passes are
 * enqueued on `stream`, and the call synchronises it twice (the host derives
 * each pass's parameters from the previous reductions); it returns with the
 * output written. */
int64_t mxb_combine_workspace_bytes(void);
int mxb_device_combine_fields(const uint16_t *dev_gauss, const uint16_t *dev_edt, int64_t nz,
                              int64_t ny, int64_t nx, const uint8_t *dev_cylinder,
                              uint16_t *dev_out, void *dev_workspace, void *stream);

/* histograms-pybind.cc:78-108 -> cpu/histograms.cc:231-306.  Adds into
 * bins[field_bins][voxel_bins] (uint64) the counts of voxel/field bin pairs:
 * voxel v (outside [vmin, vmax] -> masked) at (Z, Y, X) of the (nZ, nY, nX)
 * chunk pairs with field[floor(Z*nz/nZ) (clamped), floor(Y*ny/nY),
 * floor(X*nx/nX)]; pairs with v or the field value 0 are not counted; bin =
 * floor((bins-1) * (value - min) / (max - min)) in float64.  y/x offsets are
 * zero (the reference's callers pass (z, 0, 0)); the z offset does not enter
 * the field histogram. */
int mxb_field_histogram(const uint16_t *voxels, int64_t nZ, int64_t nY, int64_t nX,
                        const uint16_t *field, int64_t nz, int64_t ny, int64_t nx, uint64_t *bins,
                        int64_t field_bins, int64_t voxel_bins, double vmin, double vmax,
                        double fmin, double fmax, int verbose);

/* histograms-pybind.cc:16-51 -> cpu/histograms.cc:14-78.  The (nz, ny, nx)
 * chunk starts at global layer z_offset; adds voxel-bin counts into
 * x_bins[nx][vb], y_bins[ny][vb], z_bins[Nz][vb] and r_bins[Nr][vb], r =
 * floor(sqrt((x-cx)^2 + (y-cy)^2)); layers at or beyond Nz and radii at or
 * beyond Nr are skipped. */
int mxb_axis_histograms(const uint16_t *voxels, int64_t nz, int64_t ny, int64_t nx,
                        int64_t z_offset, uint64_t *x_bins, uint64_t *y_bins, uint64_t *z_bins,
                        uint64_t *r_bins, int64_t Nz, int64_t Nr, int64_t voxel_bins, uint64_t cy,
                        uint64_t cx, double vmin, double vmax, int verbose);

/* Device forms of the histograms: all pointers in HBM, enqueued on `stream`
 * (the bins are added to); each call returns after synchronising `stream`
 * (its lookup tables are staged from host memory).  The workspace holds the bin lookup tables and
 * coordinate maps; mxb_histogram_workspace_bytes gives its size for the
 * larger of the two uses with these extents. */
int64_t mxb_histogram_workspace_bytes(int64_t nZ, int64_t nY, int64_t nX);
int mxb_device_field_histogram(const uint16_t *dev_voxels, int64_t nZ, int64_t nY, int64_t nX,
                               const uint16_t *dev_field, int64_t nz, int64_t ny, int64_t nx,
                               uint64_t *dev_bins, int64_t field_bins, int64_t voxel_bins,
                               double vmin, double vmax, double fmin, double fmax,
                               void *dev_workspace, void *stream);
int mxb_device_axis_histograms(const uint16_t *dev_voxels, int64_t nz, int64_t ny, int64_t nx,
                               int64_t z_offset, uint64_t *dev_x_bins, uint64_t *dev_y_bins,
                               uint64_t *dev_z_bins, uint64_t *dev_r_bins, int64_t Nz, int64_t Nr,
                               int64_t voxel_bins, uint64_t cy, uint64_t cx, double vmin,
                               double vmax, void *dev_workspace, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MAXIBONE_FIELDS_H */
