// vsm_kernels.h -- host-side launchers of the HIP kernels (internal).
#pragma once
#include "vsm_device.h"
#include "vissat_mvs.h"

namespace vsm {

// Half-step kernel for up to maxs sources on tiles of tw x (512 / tw)
// pixels: 16 or 32 wide, and 64 for at most 16 sources.
bool tile_width_supported(int tw, int maxs);
hipError_t launch_halfstep(int maxs, int tw, const ViewDesc *views, const Tile *tiles, int ntiles,
                           const PMConst &c, Counters *counters, unsigned int *work, int grid,
                           float *c0_scratch, hipStream_t st);
// Resident workgroups of the persistent half-step kernel on this device.
int halfstep_grid(int maxs, int tw);
hipError_t launch_init(const ViewDesc *views, const Tile *tiles, int ntiles, const PMConst &c,
                       int random_hyp, int tw, hipStream_t st);
hipError_t launch_eval_cost(int maxs, const ViewDesc *views, const Tile *tiles, int ntiles,
                            const PMConst &c, float *out, hipStream_t st);
hipError_t launch_filter(int maxs, const ViewDesc *views, const Tile *tiles, int ntiles,
                         const PMConst &c, uint32_t *mask, hipStream_t st);
// Cost of the stored hypotheses against every source, sources outer, the
// window taps read from an LDS-staged box of the source (stage != 0, where the
// box fits) or from the global texture; *n_staged counts staged (tile, source)
// pairs.
hipError_t launch_eval_stage(int maxs, int stage, const ViewDesc *views, const Tile *tiles, int ntiles,
                             const PMConst &c, float *out, unsigned int *n_staged, hipStream_t st);
hipError_t launch_finalize(const ViewDesc *views, int view, int npix, float *depth, float *normal,
                           float *cost, float *depth_keep, hipStream_t st);
hipError_t launch_load_hyp(const ViewDesc *views, int view, int npix, const float *depth,
                           const float *normal, float cos_max_normal, hipStream_t st);
hipError_t launch_advance(ViewDesc *views, int nviews, hipStream_t st);
hipError_t launch_pack_depth4(const float *depth, int W, int H, float4 *out, hipStream_t st);
hipError_t launch_pack_pair(const uint8_t *gray, int W, int H, uint16_t *pair, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long *p, int n, unsigned long long v, hipStream_t st);

// Fusion (vsm_fuse.hip)
struct DsmGridDev {
    double ul_e, ul_n, e_res, n_res;
    int e_size, n_size;
};
// 3D fusion (vsm_fuse.hip)
struct Fuse3dView {
    const float *depth;             // W*H
    const float *normal;            // W*H*3, camera frame
    const uint8_t *gray;            // W*H
    int W, H;
    double P[16];                   // P4 (projection, rows 0..2 used)
    double Pinv[16];                // back-projection
    float R[9];                     // camera -> world rotation of normals
};
struct Fuse3dConst {
    int n, min_num;
    double max_depth_err, max_reproj;
    float cos_max_normal;
};
struct FusedPoint {
    double x, y, z;
    float nx, ny, nz;
    uint32_t gray;
};
hipError_t launch_fuse3d(const Fuse3dView *views, int i, int W, int H, Fuse3dConst k, FusedPoint *out,
                         int *flag, hipStream_t st);
// Compact the flagged points of one image into xyz / normal / gray at their
// running offsets; *n_sel receives the number selected (device int).
hipError_t compact_fused(const FusedPoint *pts, const int *flag, long long n, void *temp,
                         size_t &temp_bytes, FusedPoint *sel, int *n_sel, hipStream_t st);
hipError_t launch_unpack_fused(const FusedPoint *sel, long long n, double *xyz, float *normal,
                               uint8_t *gray, hipStream_t st);

// geo: NULL = grid in the ENU frame of the points; else the UTM grid.
hipError_t launch_rasterize_view(const float *depth, int W, int H, const double *invP4_dev,
                                 DsmGridDev g, const vsm_geo_frame *geo,
                                 unsigned long long *cell_max, hipStream_t st);
hipError_t launch_enu_to_utm(const double *enu, long long n, const vsm_geo_frame &geo, double *out,
                             hipStream_t st);
hipError_t launch_rasterize_points(const double *pts, long long n, DsmGridDev g,
                                   unsigned long long *cell_max, hipStream_t st);
hipError_t launch_decode_fill_f32(const unsigned long long *cell_max, int e_size, int n_size,
                                  float *dsm, hipStream_t st);
hipError_t launch_decode_fill_f64(const unsigned long long *cell_max, int e_size, int n_size,
                                  double *dsm, hipStream_t st);
hipError_t launch_median3(const float *in, int e_size, int n_size, float *out, hipStream_t st);
// Per-view DSM: decode + hole fill + 3x3 median of `keys` into `out` in one
// pass; zeroes the same cells of `clear` (a different key grid).
hipError_t launch_view_dsm(const unsigned long long *keys, unsigned long long *clear, int e_size, int n_size,
                           float *out, hipStream_t st);
hipError_t launch_fuse_stack(const float *stack, int nviews, long long ncell,
                             int min_measurements, float *out, hipStream_t st);
hipError_t launch_fill_u32(unsigned int *p, long long n, unsigned int v, hipStream_t st);
hipError_t launch_fill_u64n(unsigned long long *p, long long n, unsigned long long v,
                            hipStream_t st);

// Allocation-free proj_to_grid (vsm_fuse.hip): the cell keys live in a
// file-scope device array of kNoAllocMaxCells entries.  scratch_init takes
// its address on `device` (which must be current), once per device;
// proj_to_grid_noalloc then rasterises on stream st without allocating.
constexpr long long kNoAllocMaxCells = VSM_NOALLOC_MAX_CELLS;
hipError_t scratch_init(int device);
hipError_t proj_to_grid_noalloc(int device, const double *pts, long long n, DsmGridDev g,
                                double *dsm, hipStream_t st);

}  // namespace vsm
