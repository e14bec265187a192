// vsm_device.h -- device-side structures shared by the PatchMatch and fusion
// kernels (gfx950, wave64).  Host code sees the same layouts.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vsm {

constexpr int kMaxSrc = 32;       // sources per reference view
constexpr int kNHyp = 8;          // hypotheses per pixel per half-step
constexpr int kTileW = 32;        // tile columns (16 active pixels per row per colour)
constexpr int kTileH = 16;        // tile rows
constexpr int kThreads = 256;     // one active pixel per thread
constexpr int kMaxRadius = 8;
constexpr int kMaxSamples = 16;
constexpr int kLdsW = kTileW + 2 * kMaxRadius;
constexpr int kLdsH = kTileH + 2 * kMaxRadius;
constexpr float kMaxCost = 2.0f;

// Launch-wide constants (kernel argument).
struct PMConst {
    int radius, step, ntaps;        // ntaps per axis = 2*radius/step + 1
    float css;                      // log2(e) / (2 sigma_spatial^2)
    float ccc;                      // log2(e) / (2 (255 sigma_color)^2)   (DN units)
    float min_var;                  // 1e-5 * 255^2                        (DN^2)
    float inv2n;                    // 1 / (2 ncc_sigma^2)
    float inv2i;                    // 1 / (2 incident_angle_sigma^2)
    int num_samples;
    float geom_reg, geom_max;
    float f_min_ncc, f_min_tri, f_geom_max;
    int f_min_consistent;
    float perturbation, normal_perturbation, cos_max_normal, one_minus_cos_max, decay_base;
    int n_far;
    uint32_t seed_lo, seed_hi;
    int draw_group;                 // Monte-Carlo draws keyed by pixel (0) or by the g x 128/g
                                    // pixel group of one wave of a g-wide tile (g = draw_group)
};

// One reference view and its sources (device memory; read through scalar loads).
struct ViewDesc {
    int W, H, n_src, pass;
    int step, num_steps;            // half-step counter: iter = step >> 1, colour = step & 1
    uint32_t view_id;
    int active;
    const uint8_t *ref;             // W*H gray, row-major
    float4 *hyp;                    // (d, n0, n1, n2) per pixel
    float *cost;                    // per pixel
    float *prob;                    // n_src planes of W*H
    float depth_min, depth_max;
    float J[9];                     // d(x,y,d)/dX at the scene centre, normalised
    float N_out[9];                 // (col,row,d)-frame normal -> camera frame
    float N_in[9];                  // camera frame -> (col,row,d) frame
    const uint8_t *src_pair[kMaxSrc];   // pair texture of each source ((W + 1) x H x 2 B)
    const float4 *src_depth4[kMaxSrc];   // photometric depth of each source, 2x2 per texel
    int src_w[kMaxSrc], src_h[kMaxSrc];
    float ptri[kMaxSrc];            // triangulation prior
    float tri[kMaxSrc];             // triangulation angle (rad)
    float g[kMaxSrc][3];            // J v_s
    double M[kMaxSrc][12];          // ref (x,y,1,d) -> src homogeneous pixel
    double Mb[kMaxSrc][12];         // src (x,y,1,d) -> ref homogeneous pixel
};

// Integer anchor of one source for one tile: the texel nearest the projection
// (qx, qy, qz) of the tile centre at mid-depth, clamped into the W x H image
// (a projection behind the camera, at infinity or NaN gives a point inside
// it).  The window taps are formed relative to the anchor and, wherever a
// window may leave the image, clamped to the image box, so every relative
// texel coordinate a gather uses lies within the image side (< 2^15): the
// signed 24-bit products of the gather offsets cannot wrap.
__host__ __device__ inline void source_anchor(double qx, double qy, double qz, int W, int H, int &ax,
                                              int &ay)
{
    double px = qz > 0.0 ? qx / qz : 0.0, py = qz > 0.0 ? qy / qz : 0.0;
    px = fmin(fmax(px, 0.0), (double)(W - 1));      // fmax(NaN, 0) = 0
    py = fmin(fmax(py, 0.0), (double)(H - 1));
    ax = (int)rint(px);
    ay = (int)rint(py);
}

// Work item: one tile of one view.
struct Tile {
    int view, x0, y0, pad;
};

// Work counters (atomics only; they never feed results).
struct Counters {
    unsigned long long ncc_evals;
    unsigned long long geo_evals;
    unsigned long long tiles;
    unsigned long long work_next;   // persistent-kernel work cursor
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): counter-based, keyed by
// (seed, view id); counter = (pixel, step code, block, 0).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint32_t lo0 = 0xD2511F53u * c[0];
        const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
        const uint32_t lo1 = 0xCD9E8D57u * c[2];
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
        const uint32_t n0 = hi1 ^ c[1] ^ k0;
        const uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
}

__device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

__device__ __forceinline__ void rng_block(const PMConst &c, uint32_t view_id, uint32_t pixel,
                                          uint32_t step_code, uint32_t block, float u[4])
{
    uint32_t ctr[4] = {pixel, step_code, block, 0u};
    philox4x32_10(ctr, c.seed_lo, c.seed_hi ^ view_id);
    u[0] = u01(ctr[0]); u[1] = u01(ctr[1]); u[2] = u01(ctr[2]); u[3] = u01(ctr[3]);
}

__device__ __forceinline__ uint32_t step_code(int pass, int iter, int color)
{
    return ((uint32_t)pass << 24) | ((uint32_t)iter << 1) | (uint32_t)color;
}
__device__ __forceinline__ uint32_t init_code(int pass) { return ((uint32_t)pass << 24) | 0xFFFFFFu; }

}  // namespace vsm
