// vsm_kernels.h -- host-side launchers of the HIP kernels (internal).
#pragma once
#include "vsm_device.h"
#include "vissat_mvs.h"

namespace vsm {

hipError_t launch_halfstep(int maxs, const ViewDesc *views, const Tile *tiles, int ntiles,
                           const PMConst &c, Counters *counters, unsigned int *work, int grid,
                           hipStream_t st);
// Resident workgroups of the persistent half-step kernel on this device.
int halfstep_grid(int maxs);
hipError_t launch_init(const ViewDesc *views, const Tile *tiles, int ntiles, const PMConst &c,
                       int random_hyp, hipStream_t st);
hipError_t launch_eval_cost(int maxs, const ViewDesc *views, const Tile *tiles, int ntiles,
                            const PMConst &c, float *out, hipStream_t st);
hipError_t launch_filter(int maxs, const ViewDesc *views, const Tile *tiles, int ntiles,
                         const PMConst &c, uint32_t *mask, hipStream_t st);
hipError_t launch_finalize(const ViewDesc *views, int view, int npix, float *depth, float *normal,
                           float *cost, float *depth_keep, hipStream_t st);
hipError_t launch_load_hyp(const ViewDesc *views, int view, int npix, const float *depth,
                           const float *normal, float cos_max_normal, hipStream_t st);
hipError_t launch_advance(ViewDesc *views, int nviews, hipStream_t st);
hipError_t launch_pack_depth4(const float *depth, int W, int H, float4 *out, hipStream_t st);
hipError_t launch_pack_quad(const uint8_t *gray, int W, int H, uint32_t *quad, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long *p, int n, unsigned long long v, hipStream_t st);

// Fusion (vsm_fuse.hip)
struct DsmGridDev {
    double ul_e, ul_n, e_res, n_res;
    int e_size, n_size;
};
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
hipError_t launch_fuse_stack(const float *stack, int nviews, long long ncell,
                             int min_measurements, float *out, hipStream_t st);
hipError_t launch_fill_u32(unsigned int *p, long long n, unsigned int v, hipStream_t st);
hipError_t launch_fill_u64n(unsigned long long *p, long long n, unsigned long long v,
                            hipStream_t st);

}  // namespace vsm
