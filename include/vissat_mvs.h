/*
 * vissat_mvs.h -- C ABI of libvissat_mvs.so, the MI355X (gfx950) PatchMatch
 * MVS + 2.5D DSM fusion library.
 *
 * The reference crosses its native boundary as a CLI subprocess:
 *   colmap patch_match_stereo --workspace_path <mvs_dir> --PatchMatchStereo.<flag> <v> ...
 * (reference colmap_mvs_commands.py:40-54, spawned at lib/run_cmd.py:44), and
 * fuses depth maps in Python (aggregate_2p5d.py:45-89,
 * aggregate_2p5d_util.py:45-106, lib/proj_to_grid.py:41-81).  This header is
 * the in-process replacement: plain C types, caller-owned host buffers copied
 * in and out, device state owned by the context.
 *
 * Conventions: every call returns 0 (VSM_OK) or a negative VSM_E* code; no
 * exception crosses the ABI; vsm_last_error() gives the message of the last
 * failure on that context.  A context is used by one host thread at a time
 * (callers serialise), except vsm_submit/vsm_poll which the serving loop
 * drives from its own thread.
 */
#ifndef VISSAT_MVS_H
#define VISSAT_MVS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VSM_OK 0
#define VSM_EINVAL (-22)   /* bad argument (size, id, pointer, parameter) */
#define VSM_ENOMEM (-12)   /* allocation failed or HBM budget exceeded */
#define VSM_EDEVICE (-5)   /* HIP runtime / kernel failure */
#define VSM_ENOENT (-2)    /* unknown image id / ticket */
#define VSM_EAGAIN (-11)   /* serving: not admitted yet (over budget, queued) */

#define VSM_MAX_SRC 32     /* sources per reference view */
#define VSM_MAX_RADIUS 8   /* window_radius upper bound */

typedef struct vsm_ctx vsm_ctx;

/* PatchMatchStereo options.  Defaults (vsm_default_params) are the values of
 * the reference's photometric call, colmap_mvs_commands.py:40-48:
 * window_radius 3 (caller stereo_pipeline.py:444), min_triangulation_angle
 * 5.0, filter 0, geom_consistency 0, num_samples 15, num_iterations 12; the
 * consistency call (colmap_mvs_commands.py:60-75) sets geom_consistency 1,
 * geom_consistency_regularizer 0.3, geom_consistency_max_cost 3, filter 1,
 * filter_min_triangulation_angle 4.99, filter_min_ncc 0.1,
 * filter_geom_consistency_max_cost 1, filter_min_num_consistent 2,
 * num_iterations 1. */
typedef struct {
    int32_t window_radius;
    int32_t window_step;              /* tap stride inside the window (1 or 2) */
    float sigma_spatial;              /* <= 0: window_radius */
    float sigma_color;                /* on [0,1] intensities */
    float ncc_sigma;
    int32_t num_samples;              /* view-selection Monte-Carlo draws, <= 16 */
    int32_t num_iterations;
    float min_triangulation_angle;    /* degrees */
    float incident_angle_sigma;       /* radians */
    int32_t geom_consistency;
    float geom_consistency_regularizer;
    float geom_consistency_max_cost;  /* pixels */
    int32_t filter;
    float filter_min_ncc;
    float filter_min_triangulation_angle;   /* degrees */
    float filter_geom_consistency_max_cost; /* pixels */
    int32_t filter_min_num_consistent;
    float depth_min, depth_max;       /* > 0 overrides the per-request range */
    float perturbation;               /* depth perturbation, fraction of range */
    float normal_perturbation;        /* radians */
    float max_normal_angle;           /* degrees, cone of random normals */
    float perturb_decay;              /* per-iteration shrink of perturbations */
    int32_t n_far;                    /* propagation candidates per direction */
    uint64_t seed;
    /* Counter word of the Monte-Carlo source draws: 0 = the pixel (every
     * pixel draws independently), 1 = its group of tile_width x
     * (128 / tile_width) pixels, the pixels one wave of the half-step kernel
     * covers (32 x 4 for 32-wide tiles): each pixel still draws from its own
     * weights, with the uniforms its group shares, so neighbouring pixels
     * with similar weights draw the same sources and a wave runs fewer
     * window passes. */
    int32_t draw_group;
    /* Half-step tile width in pixels, 16, 32 or 64 (tiles of 16 x 32,
     * 32 x 16 or 64 x 8; 256 threads, one pixel of the active colour each):
     * a schedule knob of the batched form chosen by the autotuner.  Results
     * agree across widths to FP32 rounding (the per-tile FP64 re-centring of
     * the source maps sits at the tile centre). */
    int32_t tile_width;
} vsm_params;

/* One reference view to solve. */
typedef struct {
    int32_t ref_id;                   /* image id given to vsm_add_image */
    int32_t n_src;
    const int32_t *src_ids;           /* n_src image ids */
    float depth_min, depth_max;       /* depth_ranges.txt (reparam_depth.py:187-189) */
    /* Optional initial hypotheses for the geometric pass when the context
     * holds no photometric result for ref_id (resume from files,
     * use_exist_photom, colmap_mvs_commands.py:64): H*W depth and H*W*3
     * camera-frame normals, as read from the workspace. NULL otherwise. */
    const float *init_depth;
    const float *init_normal;
} vsm_view_req;

/* Caller-owned outputs; any pointer may be NULL (result stays on device). */
typedef struct {
    float *depth;                     /* H*W, row-major; 0 = no depth */
    float *normal;                    /* H*W*3, camera frame, facing the camera */
    float *cost;                      /* H*W aggregated cost */
    uint32_t *consistency;            /* H*W bitmask of consistent sources (filter) */
} vsm_view_out;

/* DSM grid in the frame of the back-projected points (ENU metres):
 * row = floor((ul_n - N)/e_res), col = floor((E - ul_e)/n_res), keeping the
 * reference's pairing of x/y resolutions (lib/proj_to_grid.py:42-43, called
 * with (e_resolution, n_resolution) at produce_dsm.py:56); the two are equal
 * in every configuration the reference uses (produce_dsm.py:41-42). */
typedef struct {
    double ul_e, ul_n;
    double e_res, n_res;
    int32_t e_size, n_size;
} vsm_dsm_grid;

/* Geodetic frame of the local ENU coordinates and UTM zone of the DSM grid.
 * The ENU origin is the AOI centre at alt_min (coordinate_system.py:41-51:
 * lat0 = (lat_min + lat_max)/2, lon0 = (lon_min + lon_max)/2, alt0 = alt_min);
 * the UTM zone and hemisphere are aoi.json's (the reference takes them from
 * the first point, lib/latlon_utm_converter.py:42-50, the same zone for an AOI
 * inside one zone).  WGS84 throughout. */
typedef struct {
    double lat0_deg, lon0_deg, alt0;
    int32_t zone_number;              /* 1..60 */
    int32_t south;                    /* 0: 'N' (false northing 0), 1: 'S' (1e7 m) */
} vsm_geo_params;

/* FP64 constants of the ENU -> geodetic -> UTM chain, filled by
 * vsm_geo_frame_init (30 doubles, no padding). */
typedef struct {
    double origin[3];                 /* ECEF of the ENU origin */
    double rot[9];                    /* ECEF offset = rot * (e, n, u), row-major */
    double cm_cos, cm_sin;            /* central meridian of the zone */
    double a, e, e2, e2a, ep2b, one_minus_f;   /* WGS84: a, e, e^2, e^2 a, e'^2 b, 1 - f */
    double k0A;                       /* 0.9996 * rectifying radius */
    double alpha[6];                  /* Krueger series coefficients (order n^6) */
    double false_easting, false_northing;
    double max_q;                     /* domain of the series (see vsm_enu_to_utm) */
} vsm_geo_frame;

/* Host-only: derive the frame constants.  Returns VSM_EINVAL for a zone
 * outside 1..60 or |lat0| > 84 degrees. */
int vsm_geo_frame_init(const vsm_geo_params *params, vsm_geo_frame *out);

/* Select the frame of the DSM grid for vsm_fuse_dsm: NULL (the default) keeps
 * the local ENU grid of the back-projected points; a frame makes every
 * back-projected point (E_enu, N_enu, U) go through local_to_global +
 * latlon_to_eastnorh (aggregate_2p5d_util.py:95-98) first, so the grid is the
 * UTM grid of aoi.json and the cell value the WGS84 altitude. */
int vsm_set_geo_frame(vsm_ctx *ctx, const vsm_geo_frame *frame);

/* Stand-alone conversion (coordinate_system.py:41-51 followed by
 * lib/latlon_utm_converter.py:39-52) of n local ENU points (n x 3 FP64) to
 * n x 3 (easting, northing, altitude) on the device, with the frame set by
 * vsm_set_geo_frame (VSM_EINVAL if none).  A point more than about 14 degrees
 * of longitude from the central meridian is outside the series' domain and
 * comes back as NaN. */
int vsm_enu_to_utm(vsm_ctx *ctx, const double *enu, int64_t n, double *out);

/* Kernel timing and work counters accumulated since the last reset. */
typedef struct {
    double halfstep_ms;               /* device time of half-step kernels */
    int64_t halfstep_launches;
    int64_t ncc_evals;                /* (pixel, hypothesis, source) NCC evaluations */
    int64_t geo_evals;
    int64_t taps;                     /* ncc_evals * window taps */
    double fuse_ms;
    int64_t fuse_launches;
    int64_t fuse_pixels;
    uint64_t hbm_bytes_in_use;
    uint64_t hbm_bytes_peak;
} vsm_stats;

void vsm_default_params(vsm_params *p);

int vsm_create(int device, uint64_t hbm_budget_bytes, vsm_ctx **out);
int vsm_destroy(vsm_ctx *ctx);
const char *vsm_last_error(vsm_ctx *ctx);
int vsm_set_params(vsm_ctx *ctx, const vsm_params *p);
int vsm_get_params(vsm_ctx *ctx, vsm_params *p);

/* Upload one grayscale image and its FP64 4x4 P4 (row-major), the
 * plane-plus-parallax projection [K[R|t]; depth_min*(0,0,1,-z0)] of
 * reparam_depth.py:136-141.  Replaces the binary reading images/ and
 * sparse/ from the workspace.  gray may be NULL to register geometry only
 * (fusion of depth maps read from files); such an image cannot be solved. */
int vsm_add_image(vsm_ctx *ctx, int32_t image_id, const uint8_t *gray, int32_t w, int32_t h,
                  int32_t row_stride, const double P4[16]);
int vsm_remove_image(vsm_ctx *ctx, int32_t image_id);

/* Provide a depth map for image_id from outside: kind 0 photometric (read
 * back from depth_maps/<img>.photometric.bin for the consistency pass,
 * use_exist_photom, colmap_mvs_commands.py:64), kind 1 geometric (for
 * fusion of maps read from files, aggregate_2p5d_util.py:75). */
int vsm_set_depth_map(vsm_ctx *ctx, int32_t image_id, int32_t kind, const float *depth, int32_t w,
                      int32_t h);

/* Provide the normal map that goes with a depth map (kind 0 or 1): w*h*3
 * floats, unit normals in the camera frame facing the camera, as in
 * normal_maps/<img>.<kind>.bin.  vsm_solve keeps its own normals the same
 * way; vsm_fuse_points reads them. */
int vsm_set_normal_map(vsm_ctx *ctx, int32_t image_id, int32_t kind, const float *normal,
                       int32_t w, int32_t h);

/* Copy the depth map of kind 0 (photometric) or 1 (geometric) the context
 * holds for image_id into a caller-owned w*h float buffer: the map a solve
 * left on the device, i.e. what depth_maps/<img>.<kind>.bin holds after the
 * reference's patch_match_stereo run (colmap_mvs_commands.py:39-82).
 * VSM_ENOENT if there is none yet. */
int vsm_get_depth_map(vsm_ctx *ctx, int32_t image_id, int32_t kind, float *depth, int32_t w,
                      int32_t h);

/* Solve n reference views in one batched launch sequence and block until
 * done.  pass 0 = photometric (run_photometric_mvs,
 * colmap_mvs_commands.py:39-55), pass 1 = geometric consistency + filter
 * (run_consistency_check, colmap_mvs_commands.py:58-82).  The geometric
 * pass starts from the request's init_depth/init_normal when given, else
 * from the reference image's photometric depth and normal maps on the device
 * (a photometric solve keeps them; vsm_set_depth_map/vsm_set_normal_map set
 * them from files).  With use_graph != 0 the half-step sequence is captured
 * once in a hipGraph and replayed. */
int vsm_solve(vsm_ctx *ctx, const vsm_view_req *reqs, int32_t n, int32_t pass,
              vsm_view_out *outs, int32_t use_graph);

/* vsm_solve in stages (the same launches and results): vsm_batch_begin
 * prepares the n views and sets their initial hypotheses (*batch receives a
 * handle); vsm_batch_step runs up to n_halfsteps more red-black half-steps
 * over all of them, one persistent launch per half-step (with use_graph the
 * sequence of that length is captured into a hipGraph on first use and
 * replayed on later calls), and sets *remaining (may be NULL) to the
 * half-steps left of the pass's 2 * num_iterations; vsm_batch_end runs the
 * filter (if the parameters at begin ask for it), writes the outputs as
 * vsm_solve does and releases the batch -- VSM_EINVAL while half-steps
 * remain, unless discard != 0, which releases it without outputs.  The
 * parameters in force at begin hold for the whole batch.  Lets a caller time
 * or interleave the half-steps of several solves; the reference runs the
 * whole pass inside one subprocess (colmap_mvs_commands.py:39-82). */
int vsm_batch_begin(vsm_ctx *ctx, const vsm_view_req *reqs, int32_t n, int32_t pass,
                    int32_t use_graph, int32_t *batch);
int vsm_batch_step(vsm_ctx *ctx, int32_t batch, int32_t n_halfsteps, int32_t *remaining);
int vsm_batch_end(vsm_ctx *ctx, int32_t batch, vsm_view_out *outs, int32_t discard);

/* Serving: non-blocking admission of per-view solves (the request is copied).
 * vsm_submit_pass queues a solve of pass 0 (photometric,
 * colmap_mvs_commands.py:39-55) or 1 (geometric consistency,
 * colmap_mvs_commands.py:58-82), with the filter on (1), off (0) or as
 * params.filter (-1); vsm_submit is vsm_submit_pass(pass 0, filter -1).
 * Results are those of solving the requests one at a time in submission
 * order: a geometric request starts once every earlier photometric request of
 * its reference or its sources has finished (and fails with VSM_EINVAL when
 * the reference has no photometric result or a source no photometric depth
 * map and no such request is pending); a photometric request waits for every
 * earlier request that reads or writes its image's photometric maps.  A
 * request that can start is admitted when its HBM footprint fits the budget
 * (ready requests are admitted in submission order; one that cannot fit even
 * with nothing else active fails with VSM_ENOMEM), joins the running batch at
 * the next half-step boundary and shares its launches with the others; its
 * result does not depend on which requests share them.  Parameters in force
 * at each vsm_serve_step apply to every active request: change them only
 * between requests.  vsm_poll copies the result into out when done
 * (*done = 1), else *done = 0.  A finished request's depth and normals also
 * become its image's device maps of that pass (vsm_get_depth_map,
 * vsm_fuse_dsm; a geometric pass starts from the photometric ones), as after
 * vsm_solve.
 * vsm_serve_step admits what can start and advances every admitted request
 * by one half-step (one batched launch).  An image used by a queued or
 * active request cannot be removed or replaced. */
int vsm_submit(vsm_ctx *ctx, const vsm_view_req *req, uint64_t *ticket);
int vsm_submit_pass(vsm_ctx *ctx, const vsm_view_req *req, int32_t pass, int32_t filter,
                    uint64_t *ticket);
int vsm_serve_step(vsm_ctx *ctx, int32_t *n_active, int32_t *n_queued);
int vsm_poll(vsm_ctx *ctx, uint64_t ticket, vsm_view_out *out, int32_t *done);

/* Per-pixel cost of fixed hypotheses against every source: out[s*H*W + p].
 * depth H*W, normal H*W*3 in the (col,row,d) frame used by the solver.
 * pass 1 adds the geometric term.  Test entry point for the
 * "cost under fixed hypotheses" parity check. */
int vsm_eval_cost(vsm_ctx *ctx, const vsm_view_req *req, int32_t pass, const float *depth,
                  const float *normal_uvd, float *out);

/* Measurement entry point (no counterpart in the reference): the same costs
 * as vsm_eval_cost, sources outer, with the window taps read either from the
 * source texture through the caches (stage 0) or from a box of the source
 * staged in LDS by each workgroup for each source (stage 1; a box of more than
 * 4096 texels falls back to the caches).  Runs the kernel `repeats` times and
 * reports the average kernel time and the fraction of (tile, source) pairs
 * that were staged; out as in vsm_eval_cost (identical for both stages). */
int vsm_eval_cost_staging(vsm_ctx *ctx, const vsm_view_req *req, int32_t pass, const float *depth,
                          const float *normal_uvd, int32_t stage, int32_t repeats, float *out,
                          float *ms_per_launch, float *staged_fraction);

/* One red-black half-step on host-provided state (in/out): depth H*W,
 * normal_uvd H*W*3, cost H*W, prob n_src*H*W.  Test entry point for
 * lockstep parity against the oracle. */
int vsm_debug_halfstep(vsm_ctx *ctx, const vsm_view_req *req, int32_t pass, int32_t iter,
                       int32_t color, float *depth, float *normal_uvd, float *cost, float *prob);
/* Random initialisation of the state (depth, normal_uvd, cost, prob). */
int vsm_debug_init(vsm_ctx *ctx, const vsm_view_req *req, int32_t pass, float *depth,
                   float *normal_uvd, float *cost, float *prob);
/* Filter on host-provided state; writes depth/normal_uvd in place and the mask. */
int vsm_debug_filter(vsm_ctx *ctx, const vsm_view_req *req, int32_t pass, float *depth,
                     float *normal_uvd, uint32_t *mask);

/* 2.5D fusion of the depth maps of n images held by the context into the DSM
 * grid: per view back-projection through inv(P4) (aggregate_2p5d_util.py:86-91),
 * per-cell nanmax + 3x3 hole fill (lib/proj_to_grid.py:41-81), 3x3 median
 * (produce_dsm.py:58); across views the <= (min_measurements-1) rejection,
 * MAD outlier removal, mean and 3x3 median (aggregate_2p5d.py:66-81).
 * depth_kind: 0 photometric, 1 geometric (aggregate_2p5d.py:48).
 * With a geo frame set (vsm_set_geo_frame) each back-projected ENU point is
 * converted to UTM (E, N) and WGS84 altitude before it is rasterised.
 * inv_proj_mats: optional n*16 FP64 row-major inverse projections as read
 * from inv_proj_mats.txt (aggregate_2p5d_util.py:55-61); NULL uses the
 * context's own inverse of each P4.
 * dsm_out: n_size*e_size float (NaN = empty); per_view_out: optional
 * n*n_size*e_size per-view DSMs. */
int vsm_fuse_dsm(vsm_ctx *ctx, const vsm_dsm_grid *grid, const int32_t *image_ids, int32_t n,
                 int32_t depth_kind, int32_t min_measurements, const double *inv_proj_mats,
                 float *dsm_out, float *per_view_out);

/* Stand-alone rasteriser with the semantics of proj_to_grid
 * (lib/proj_to_grid.py:41-81): points n x 3 FP64 (E, N, h) -> n_size x
 * e_size FP64 grid with per-cell nanmax and the 3x3 hole fill. */
int vsm_proj_to_grid(vsm_ctx *ctx, const double *points, int64_t n, const vsm_dsm_grid *grid,
                     double *dsm_out);

/* proj_to_grid without allocating (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  Same semantics and
 * bit-identical output as vsm_proj_to_grid, but points_dev (n x 3 FP64) and
 * dsm_dev (n_size x e_size FP64) are device pointers, the grid may hold at
 * most VSM_NOALLOC_MAX_CELLS cells (VSM_EINVAL beyond), and the work is
 * queued on `stream` (a hipStream_t; NULL = the context's stream) without
 * waiting: the result is ready when that stream is (vsm_synchronize for the
 * context's stream).  The per-cell scratch is one device array per GPU,
 * allocated with the library's code and held until the process ends, shared
 * by every call on that GPU: calls that overlap on one GPU race. */
#define VSM_NOALLOC_MAX_CELLS (4ll * 1024 * 1024)   /* 32 MiB of 64-bit cell keys */
int vsm_proj_to_grid_noalloc(vsm_ctx *ctx, const double *points_dev, int64_t n,
                             const vsm_dsm_grid *grid, double *dsm_dev, void *stream);

/* 3D fusion of depth maps into a point cloud, the path of aggregate_3d.py
 * (aggregate_3d.py:44-82: `colmap stereo_fusion --input_type geometric
 * --StereoFusion.min_num_pixels 4 --StereoFusion.max_reproj_error 2
 * --StereoFusion.max_depth_error 1.0 --StereoFusion.max_normal_error 10`),
 * order-independent: a pixel p of image i with depth d > 0 back-projects to
 * X; image j is consistent with it when X projects inside j onto pixel q
 * (nearest) with depth d_j > 0, the back-projection Y of q lies within
 * max_depth_error metres of X, Y re-projects into i within
 * max_reproj_error pixels of p, and the two world-frame normals differ by at
 * most max_normal_error degrees.  p yields a point when at least
 * min_num_pixels pixels (p included) agree and no consistent image has a
 * smaller index than i (each cluster is emitted once, by its first image):
 * the mean of X and the Y's, the normalised sum of the normals and the
 * rounded mean grey value.  Points come out in image order, then row-major
 * pixel order. */
typedef struct {
    int32_t min_num_pixels;           /* 4 */
    float max_reproj_error;           /* 2 px */
    float max_depth_error;            /* 1.0 m */
    float max_normal_error_deg;       /* 10 degrees */
} vsm_fuse3d_params;

void vsm_default_fuse3d_params(vsm_fuse3d_params *p);

/* image_ids: n images with depth (kind) and normal maps held by the context.
 * inv_proj_mats: optional n*16 inverse projections (NULL: the context's);
 * cam_to_world: n*9 row-major rotations taking camera-frame normals to the
 * world (ENU) frame (R^T of P4 = [K R | K t; ...]).  Writes up to capacity
 * points (xyz n x 3 FP64, normal n x 3 float or NULL, grey n uint8 or NULL)
 * and sets *count to the total, which may exceed capacity (nothing beyond
 * capacity is written; call again with a larger buffer). */
int vsm_fuse_points(vsm_ctx *ctx, const int32_t *image_ids, int32_t n, int32_t depth_kind,
                    const vsm_fuse3d_params *params, const double *inv_proj_mats,
                    const double *cam_to_world, int64_t capacity, double *xyz, float *normal,
                    uint8_t *grey, int64_t *count);

/* 3x3 median with replicated border on an n_size x e_size float raster
 * (cv2.medianBlur(dsm, 3) at produce_dsm.py:58 and aggregate_2p5d.py:81);
 * NaN excluded from each window, all-NaN -> NaN. */
int vsm_median_blur3(vsm_ctx *ctx, const float *in, int32_t e_size, int32_t n_size, float *out);

/* Change the context's HBM budget (0 = unlimited); freed blocks the context
 * caches beyond it are released.  Later allocations and serving admissions
 * are checked against the new value; nothing already allocated is moved.
 * The reference has no budget (the colmap binary takes what it needs). */
int vsm_set_hbm_budget(vsm_ctx *ctx, uint64_t hbm_budget_bytes);

int vsm_get_stats(vsm_ctx *ctx, vsm_stats *out);
int vsm_reset_stats(vsm_ctx *ctx);
/* Enable (1) or disable (0) per-launch event timing of half-step kernels. */
int vsm_set_timing(vsm_ctx *ctx, int32_t enabled);
int vsm_synchronize(vsm_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* VISSAT_MVS_H */
