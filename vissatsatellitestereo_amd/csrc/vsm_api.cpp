// vsm_api.cpp -- C ABI of libvissat_mvs.so (declared in include/vissat_mvs.h).
//
// Host side of the MVS path: context and HBM budget, image upload, FP64
// geometry setup per reference view, batched red-black PatchMatch solves
// (optionally replayed from a captured hipGraph), filter and outputs, the
// serving loop (continuous batching with HBM-footprint admission) and 2.5D
// DSM fusion.  Replaces the `colmap patch_match_stereo` subprocess of the
// reference (colmap_mvs_commands.py:39-82, lib/run_cmd.py:44) and the Python
// fusion (aggregate_2p5d.py:45-89).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "vissat_mvs.h"
#include "vsm_device.h"
#include "vsm_kernels.h"

using namespace vsm;

namespace {

struct ImageRec {
    int id = 0, W = 0, H = 0;
    uint8_t *gray = nullptr;
    uint16_t *pair = nullptr;               // pair texture, (W + 1) x H texels
    double P4[16];
    double Pinv[16];
    float *depth[2] = {nullptr, nullptr};   // photometric / geometric depth kept on device
    float *normal[2] = {nullptr, nullptr};  // camera-frame normals that go with depth[]
    float4 *depth4 = nullptr;               // photometric depth, 2x2 neighbourhood per texel
    uint64_t depth_ver = 0, depth4_ver = 0; // depth4 is current when the versions agree
};

struct ViewState {
    ViewDesc desc;                          // host copy
    float4 *hyp = nullptr;
    float *cost = nullptr;
    float *prob = nullptr;
};

// A staged batched solve (vsm_batch_begin / _step / _end): the views' state,
// the device view descriptors and tile list, and the instantiated hipGraphs
// of its half-step sequences, kept for reuse across step calls.
struct Batch {
    int pass = 0;
    bool use_graph = false;
    vsm_params params;                      // snapshot at begin: every step and the filter use it
    std::vector<ViewState> views;
    std::vector<int> ref_ids;
    ViewDesc *dviews = nullptr;
    Tile *dtiles = nullptr;
    int ntiles = 0, maxs = 1;
    int steps_done = 0, nsteps = 0;
    std::vector<std::pair<int, hipGraphExec_t>> graphs;   // (half-steps per replay, executable)
};

struct ServeReq {
    uint64_t ticket = 0;
    int ref = 0;
    int pass = 0;                           // 0 photometric, 1 geometric consistency
    int filter = 0;
    std::vector<int> srcs;
    float dmin = 0.f, dmax = 0.f;
    // device state while active
    ViewState vs;
    int host_step = 0;
    int slot = -1;                          // its descriptor in the served view array
    Tile *dtiles = nullptr;                 // its half-step tiles (view index = slot)
    std::vector<Tile> htiles;               // their host copy (source of the upload)
    int ntiles = 0, tile_w = 0;
    // results (device) when done: depth, normal (3), cost, consistency mask
    float *out_depth = nullptr, *out_normal = nullptr, *out_cost = nullptr;
    uint32_t *out_mask = nullptr;
    // the image's depth and normal maps of this pass when it has none yet
    // (allocated at admission, handed to the image at retirement)
    float *keep_depth = nullptr, *keep_normal = nullptr;
    bool done = false;
    int error = 0;
    std::string errmsg;
};

}  // namespace

struct vsm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0, in_use = 0, peak = 0;
    vsm_params params;
    std::map<int, ImageRec> images;
    std::map<void *, size_t> allocs;       // live device blocks
    std::multimap<size_t, void *> cache;   // freed blocks kept for reuse (by size)
    uint64_t cached_bytes = 0;
    std::string err;
    Counters *d_counters = nullptr;
    unsigned int *d_work = nullptr;        // per-XCD tile cursors of the persistent kernel
    float *d_c0 = nullptr;                 // per-workgroup view-selection costs (> 16 sources)
    size_t c0_floats = 0;
    int grid[9] = {0};                     // resident workgroups per (MAXS, tile width) variant
    bool timing = false;
    double halfstep_ms = 0.0;
    int64_t halfstep_launches = 0;
    double fuse_ms = 0.0;
    int64_t fuse_launches = 0, fuse_pixels = 0;
    bool has_geo = false;                  // DSM grid in UTM (vsm_set_geo_frame)
    vsm_geo_frame geo{};
    int64_t taps_per_eval = 0;
    int64_t acc_ncc = 0, acc_geo = 0, acc_taps = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // serving
    std::mutex serve_mu;
    uint64_t next_ticket = 1;
    std::deque<std::unique_ptr<ServeReq>> queue;
    std::vector<ServeReq *> active;
    std::map<uint64_t, std::unique_ptr<ServeReq>> reqs;
    ViewDesc *serve_views = nullptr;
    std::vector<ServeReq *> serve_slots;   // request per descriptor slot (null: free)
    Tile *serve_tiles = nullptr;           // the active requests' tiles, concatenated
    int serve_cap_tiles = 0, serve_ntiles = 0;
    bool serve_dirty = true;               // serve_tiles does not match the active list
    // staged batched solves
    std::map<int, std::unique_ptr<Batch>> batches;
    int next_batch = 1;
};

namespace {

constexpr int kServeMaxViews = 64;

int fail(vsm_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(ctx, expr)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail((ctx), VSM_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                    \
    } while (0)

// Device memory: blocks freed by the library go to a per-context cache and
// are reused by later allocations of up to the same size (all work runs in
// order on the context's one stream, so a reused block is never touched by
// work queued before its reuse).  Live blocks count against the HBM budget;
// cached blocks are released (hipFree) whenever live + cached + request would
// exceed it, when hipMalloc fails, and at vsm_destroy.
constexpr size_t kAllocAlign = 256;
constexpr uint64_t kCacheMaxBytes = 32ull << 30;

void release_cached(vsm_ctx *c, uint64_t need)
{
    while (!c->cache.empty() && (need == UINT64_MAX ||
                                 c->in_use + c->cached_bytes + need > c->budget)) {
        auto it = std::prev(c->cache.end());          // largest first
        (void)hipFree(it->second);
        c->cached_bytes -= it->first;
        c->cache.erase(it);
    }
}

int dmalloc(vsm_ctx *c, void **p, size_t bytes)
{
    *p = nullptr;
    bytes = (std::max<size_t>(bytes, 1) + kAllocAlign - 1) / kAllocAlign * kAllocAlign;
    if (c->budget && c->in_use + bytes > c->budget)
        return fail(c, VSM_ENOMEM, "HBM budget exceeded: %llu + %llu > %llu bytes",
                    (unsigned long long)c->in_use, (unsigned long long)bytes,
                    (unsigned long long)c->budget);
    auto hit = c->cache.lower_bound(bytes);
    if (hit != c->cache.end() && hit->first <= 2 * bytes) {
        *p = hit->second;
        const size_t sz = hit->first;
        c->cached_bytes -= sz;
        c->cache.erase(hit);
        c->allocs[*p] = sz;
        c->in_use += sz;
        c->peak = std::max(c->peak, c->in_use);
        return VSM_OK;
    }
    if (c->budget) release_cached(c, bytes);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess && !c->cache.empty()) {
        (void)hipGetLastError();
        release_cached(c, UINT64_MAX);
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, VSM_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    c->allocs[*p] = bytes;
    c->in_use += bytes;
    c->peak = std::max(c->peak, c->in_use);
    return VSM_OK;
}

void dfree(vsm_ctx *c, void *p)
{
    if (!p) return;
    auto it = c->allocs.find(p);
    if (it == c->allocs.end()) {
        (void)hipFree(p);
        return;
    }
    const size_t sz = it->second;
    c->in_use -= sz;
    c->allocs.erase(it);
    c->cache.emplace(sz, p);
    c->cached_bytes += sz;
    while (c->cached_bytes > kCacheMaxBytes && !c->cache.empty()) {
        auto big = std::prev(c->cache.end());
        (void)hipFree(big->second);
        c->cached_bytes -= big->first;
        c->cache.erase(big);
    }
}

template <typename T>
int dalloc(vsm_ctx *c, T **p, size_t count)
{
    return dmalloc(c, reinterpret_cast<void **>(p), count * sizeof(T));
}

// ---------------------------------------------------------------------------
// FP64 linear algebra for the geometry setup
// ---------------------------------------------------------------------------
bool inv4(const double *a, double *out)
{
    double m[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            m[i][j] = a[4 * i + j];
            m[i][4 + j] = i == j ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(m[r][c]) > std::fabs(m[piv][c])) piv = r;
        if (m[piv][c] == 0.0) return false;
        if (piv != c)
            for (int j = 0; j < 8; ++j) std::swap(m[c][j], m[piv][j]);
        const double d = m[c][c];
        for (int j = 0; j < 8; ++j) m[c][j] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const double f = m[r][c];
            if (f == 0.0) continue;
            for (int j = 0; j < 8; ++j) m[r][j] -= f * m[c][j];
        }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out[4 * i + j] = m[i][4 + j];
    return true;
}

bool inv3(const double *a, double *o)
{
    const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                       a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (det == 0.0 || !std::isfinite(det)) return false;
    const double id = 1.0 / det;
    o[0] = (a[4] * a[8] - a[5] * a[7]) * id;
    o[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    o[3] = (a[5] * a[6] - a[3] * a[8]) * id;
    o[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    o[6] = (a[3] * a[7] - a[4] * a[6]) * id;
    o[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    return true;
}

// (3x4 rows of A) x (4x4 B), normalised to max |.| = 1.
void mul34_44(const double *A, const double *B, double *out)
{
    double mx = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += A[4 * i + k] * B[4 * k + j];
            out[4 * i + j] = s;
            mx = std::max(mx, std::fabs(s));
        }
    if (mx > 0.0)
        for (int k = 0; k < 12; ++k) out[k] /= mx;
}

void camera_center(const double *P4, double C[3])
{
    double A[9] = {P4[0], P4[1], P4[2], P4[4], P4[5], P4[6], P4[8], P4[9], P4[10]}, Ai[9];
    inv3(A, Ai);
    for (int i = 0; i < 3; ++i)
        C[i] = -(Ai[3 * i] * P4[3] + Ai[3 * i + 1] * P4[7] + Ai[3 * i + 2] * P4[11]);
}

double norm3(const double *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// Per-view geometry: M_s, Mb_s, triangulation angles/priors, g_s = J v_s,
// J = d(x,y,d)/dX at the scene centre X0, and the normal frame maps.
int build_geometry(vsm_ctx *c, const ImageRec &ref, const std::vector<const ImageRec *> &srcs,
                   float dmin, float dmax, ViewDesc &V)
{
    const double dmid = 0.5 * ((double)dmin + (double)dmax);
    const double p[4] = {(ref.W - 1) / 2.0, (ref.H - 1) / 2.0, 1.0, dmid};
    double Xh[4];
    for (int i = 0; i < 4; ++i)
        Xh[i] = ref.Pinv[4 * i] * p[0] + ref.Pinv[4 * i + 1] * p[1] + ref.Pinv[4 * i + 2] * p[2] +
                ref.Pinv[4 * i + 3] * p[3];
    if (!(std::fabs(Xh[3]) > 0.0)) return fail(c, VSM_EINVAL, "degenerate P4 for image %d", ref.id);
    const double X0[3] = {Xh[0] / Xh[3], Xh[1] / Xh[3], Xh[2] / Xh[3]};
    double y[4];
    for (int i = 0; i < 4; ++i)
        y[i] = ref.P4[4 * i] * X0[0] + ref.P4[4 * i + 1] * X0[1] + ref.P4[4 * i + 2] * X0[2] +
               ref.P4[4 * i + 3];
    double J[9];
    const int rows[3] = {0, 1, 3};
    double mx = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) {
            const int i = rows[r];
            J[3 * r + k] = (ref.P4[4 * i + k] * y[2] - y[i] * ref.P4[8 + k]) / (y[2] * y[2]);
            mx = std::max(mx, std::fabs(J[3 * r + k]));
        }
    for (int k = 0; k < 9; ++k) J[k] /= mx;
    double Cr[3];
    camera_center(ref.P4, Cr);
    const double amin = (double)c->params.min_triangulation_angle * M_PI / 180.0;
    for (size_t s = 0; s < srcs.size(); ++s) {
        const ImageRec &S = *srcs[s];
        double M[12], Mb[12];
        mul34_44(S.P4, ref.Pinv, M);
        mul34_44(ref.P4, S.Pinv, Mb);
        for (int k = 0; k < 12; ++k) { V.M[s][k] = M[k]; V.Mb[s][k] = Mb[k]; }
        double Cs[3];
        camera_center(S.P4, Cs);
        const double a[3] = {Cr[0] - X0[0], Cr[1] - X0[1], Cr[2] - X0[2]};
        const double b[3] = {Cs[0] - X0[0], Cs[1] - X0[1], Cs[2] - X0[2]};
        const double na = norm3(a), nb = norm3(b);
        double cs = (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / (na * nb);
        cs = std::min(1.0, std::max(-1.0, cs));
        const double tri = std::acos(cs);
        const double tm = std::min(tri, amin);
        V.tri[s] = (float)tri;
        V.ptri[s] = (float)(amin > 0.0 ? 1.0 - (tm - amin) * (tm - amin) / (amin * amin) : 1.0);
        const double v[3] = {b[0] / nb, b[1] / nb, b[2] / nb};
        for (int r = 0; r < 3; ++r)
            V.g[s][r] = (float)(J[3 * r] * v[0] + J[3 * r + 1] * v[1] + J[3 * r + 2] * v[2]);
        V.src_w[s] = S.W;
        V.src_h[s] = S.H;
        V.src_pair[s] = (const uint8_t *)S.pair;
        V.src_depth4[s] = nullptr;
    }
    // Rotation of the reference camera from P[:3,:3] = K R.
    const double *A = ref.P4;
    double r3[3] = {A[8], A[9], A[10]};
    double n3 = norm3(r3);
    for (double &x : r3) x /= n3;
    double r2[3] = {A[4], A[5], A[6]};
    const double d23 = r2[0] * r3[0] + r2[1] * r3[1] + r2[2] * r3[2];
    for (int k = 0; k < 3; ++k) r2[k] -= d23 * r3[k];
    const double n2 = norm3(r2);
    for (double &x : r2) x /= n2;
    double r1[3] = {r2[1] * r3[2] - r2[2] * r3[1], r2[2] * r3[0] - r2[0] * r3[2],
                    r2[0] * r3[1] - r2[1] * r3[0]};
    if (r1[0] * A[0] + r1[1] * A[1] + r1[2] * A[2] < 0)
        for (double &x : r1) x = -x;
    const double R[9] = {r1[0], r1[1], r1[2], r2[0], r2[1], r2[2], r3[0], r3[1], r3[2]};
    double Nout[9], Nin[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            Nout[3 * i + j] = R[3 * i] * J[3 * j] + R[3 * i + 1] * J[3 * j + 1] + R[3 * i + 2] * J[3 * j + 2];
    if (!inv3(Nout, Nin)) return fail(c, VSM_EINVAL, "singular normal frame for image %d", ref.id);
    for (int k = 0; k < 9; ++k) {
        V.J[k] = (float)J[k];
        V.N_out[k] = (float)Nout[k];
        V.N_in[k] = (float)Nin[k];
    }
    V.depth_min = dmin;
    V.depth_max = dmax;
    return VSM_OK;
}

PMConst make_const(const vsm_params &p)
{
    PMConst c;
    c.radius = p.window_radius;
    c.step = p.window_step;
    c.ntaps = 2 * p.window_radius / p.window_step + 1;
    const double ss = p.sigma_spatial > 0 ? p.sigma_spatial : (double)p.window_radius;
    const double log2e = 1.4426950408889634;
    c.css = (float)(log2e / (2.0 * ss * ss));
    const double sc = 255.0 * p.sigma_color;
    c.ccc = (float)(log2e / (2.0 * sc * sc));
    c.min_var = (float)(1e-5 * 255.0 * 255.0);
    c.inv2n = (float)(1.0 / (2.0 * p.ncc_sigma * p.ncc_sigma));
    c.inv2i = (float)(1.0 / (2.0 * p.incident_angle_sigma * p.incident_angle_sigma));
    c.num_samples = p.num_samples;
    c.geom_reg = p.geom_consistency_regularizer;
    c.geom_max = p.geom_consistency_max_cost;
    c.f_min_ncc = p.filter_min_ncc;
    c.f_min_tri = (float)(p.filter_min_triangulation_angle * M_PI / 180.0);
    c.f_geom_max = p.filter_geom_consistency_max_cost;
    c.f_min_consistent = p.filter_min_num_consistent;
    c.perturbation = p.perturbation;
    c.normal_perturbation = p.normal_perturbation;
    const double cm = std::cos((double)p.max_normal_angle * M_PI / 180.0);
    c.cos_max_normal = (float)cm;
    c.one_minus_cos_max = (float)(1.0 - cm);
    c.decay_base = p.perturb_decay;
    c.n_far = p.n_far;
    c.seed_lo = (uint32_t)p.seed;
    c.seed_hi = (uint32_t)(p.seed >> 32);
    c.draw_group = p.draw_group ? p.tile_width : 0;     // group width in pixels, or per pixel
    return c;
}

int check_params(vsm_ctx *c, const vsm_params &p)
{
    if (p.window_radius < 1 || p.window_radius > VSM_MAX_RADIUS)
        return fail(c, VSM_EINVAL, "window_radius %d outside [1, %d]", p.window_radius, VSM_MAX_RADIUS);
    if (p.window_step < 1 || p.window_step > p.window_radius)
        return fail(c, VSM_EINVAL, "window_step %d outside [1, window_radius]", p.window_step);
    if (p.num_samples < 1 || p.num_samples > kMaxSamples)
        return fail(c, VSM_EINVAL, "num_samples %d outside [1, %d]", p.num_samples, kMaxSamples);
    if (p.num_iterations < 0 || p.num_iterations > 1000)
        return fail(c, VSM_EINVAL, "num_iterations %d outside [0, 1000]", p.num_iterations);
    if (!(p.sigma_color > 0) || !(p.ncc_sigma > 0) || !(p.incident_angle_sigma > 0))
        return fail(c, VSM_EINVAL, "sigma_color, ncc_sigma and incident_angle_sigma must be > 0");
    if (p.n_far < 1 || p.n_far > 64) return fail(c, VSM_EINVAL, "n_far %d outside [1, 64]", p.n_far);
    if (p.draw_group != 0 && p.draw_group != 1) return fail(c, VSM_EINVAL, "draw_group must be 0 or 1");
    if (!tile_width_supported(p.tile_width, 1))
        return fail(c, VSM_EINVAL, "tile_width %d: the half-step kernel has tiles 16, 32 or 64 wide", p.tile_width);
    if (!(p.max_normal_angle > 0 && p.max_normal_angle < 90))
        return fail(c, VSM_EINVAL, "max_normal_angle must be in (0, 90) degrees");
    return VSM_OK;
}

// Tiles are listed in blocks of 8 x 8 tiles (256 x 128 pixels), blocks row-major.
// The persistent half-step kernel gives each XCD a contiguous share of the list and
// its 64 resident workgroups take consecutive tiles, so at any moment they cover one
// square block, whose footprint in each source image is smaller than that of a
// 64-tile strip: +5.6 % on 20 sources of 4096^2, +1 % on 8 of 1024^2 (8 x 16,
// 16 x 8 and 16 x 16 blocks measured the same).  Tiles are independent within a
// half-step and the random numbers are keyed by view and pixel, so the order never
// changes a result.
constexpr int kBlockTiles = 8;

// Tiles of width tile_w and height 2 kThreads / tile_w (the half-step kernel's
// shape; the filter and cost kernels use kTileW x kTileH).
void tiles_for(const ViewDesc &V, int view_index, std::vector<Tile> &out, int tile_w = kTileW)
{
    const int tile_h = 2 * kThreads / tile_w;
    const int tw = (V.W + tile_w - 1) / tile_w, th = (V.H + tile_h - 1) / tile_h;
    for (int by = 0; by < th; by += kBlockTiles)
        for (int bx = 0; bx < tw; bx += kBlockTiles)
            for (int ty = by; ty < std::min(th, by + kBlockTiles); ++ty)
                for (int tx = bx; tx < std::min(tw, bx + kBlockTiles); ++tx)
                    out.push_back(Tile{view_index, tx * tile_w, ty * tile_h, 0});
}

size_t state_bytes(int W, int H, int S)
{
    const size_t hw = (size_t)W * H;
    return hw * sizeof(float4) + hw * sizeof(float) + hw * S * sizeof(float);
}

// Prepare one view: validation, geometry, state allocation.
int prepare_view(vsm_ctx *c, const vsm_view_req &r, int pass, uint32_t view_id, ViewState &vs)
{
    auto it = c->images.find(r.ref_id);
    if (it == c->images.end()) return fail(c, VSM_ENOENT, "unknown reference image %d", r.ref_id);
    const ImageRec &ref = it->second;
    if (!ref.pair) return fail(c, VSM_EINVAL, "image %d was registered without pixels", r.ref_id);
    if (r.n_src < 1 || r.n_src > VSM_MAX_SRC)
        return fail(c, VSM_EINVAL, "n_src %d outside [1, %d]", r.n_src, VSM_MAX_SRC);
    if (!r.src_ids) return fail(c, VSM_EINVAL, "src_ids is NULL");
    std::vector<ImageRec *> srcs;
    for (int s = 0; s < r.n_src; ++s) {
        auto si = c->images.find(r.src_ids[s]);
        if (si == c->images.end()) return fail(c, VSM_ENOENT, "unknown source image %d", r.src_ids[s]);
        if (r.src_ids[s] == r.ref_id) return fail(c, VSM_EINVAL, "source %d is the reference", r.ref_id);
        if (!si->second.pair) return fail(c, VSM_EINVAL, "image %d was registered without pixels", r.src_ids[s]);
        srcs.push_back(&si->second);
    }
    float dmin = c->params.depth_min > 0 ? c->params.depth_min : r.depth_min;
    float dmax = c->params.depth_max > 0 ? c->params.depth_max : r.depth_max;
    if (!(dmax > dmin) || !std::isfinite(dmin) || !std::isfinite(dmax))
        return fail(c, VSM_EINVAL, "invalid depth range [%g, %g] for image %d", dmin, dmax, r.ref_id);
    memset(&vs.desc, 0, sizeof(ViewDesc));
    ViewDesc &V = vs.desc;
    int rc = build_geometry(c, ref, std::vector<const ImageRec *>(srcs.begin(), srcs.end()), dmin,
                            dmax, V);
    if (rc) return rc;
    V.W = ref.W;
    V.H = ref.H;
    V.n_src = r.n_src;
    V.pass = pass;
    V.step = 0;
    V.num_steps = 2 * c->params.num_iterations;
    V.view_id = view_id;
    V.active = 1;
    V.ref = ref.gray;
    if (pass == 1) {
        for (int s = 0; s < r.n_src; ++s) {
            ImageRec &S = *srcs[s];
            if (!S.depth[0])
                return fail(c, VSM_EINVAL,
                            "geometric pass needs the photometric depth map of source image %d",
                            S.id);
            if (!S.depth4 || S.depth4_ver != S.depth_ver) {
                if (!S.depth4 && (rc = dalloc(c, &S.depth4, (size_t)S.W * S.H))) return rc;
                HIPCHK(c, launch_pack_depth4(S.depth[0], S.W, S.H, S.depth4, c->stream));
                S.depth4_ver = S.depth_ver;
            }
            V.src_depth4[s] = S.depth4;
        }
    }
    const size_t hw = (size_t)ref.W * ref.H;
    rc = dalloc(c, &vs.hyp, hw);
    if (!rc) rc = dalloc(c, &vs.cost, hw);
    if (!rc) rc = dalloc(c, &vs.prob, hw * r.n_src);
    if (rc) {
        dfree(c, vs.hyp); dfree(c, vs.cost); dfree(c, vs.prob);
        vs.hyp = nullptr; vs.cost = nullptr; vs.prob = nullptr;
        return rc;
    }
    V.hyp = vs.hyp;
    V.cost = vs.cost;
    V.prob = vs.prob;
    return VSM_OK;
}

void release_view(vsm_ctx *c, ViewState &vs)
{
    dfree(c, vs.hyp);
    dfree(c, vs.cost);
    dfree(c, vs.prob);
    vs.hyp = nullptr; vs.cost = nullptr; vs.prob = nullptr;
}

int grid_for(vsm_ctx *c, int maxs, int tw)
{
    const int k = (maxs <= 8 ? 0 : (maxs <= 16 ? 1 : 2)) * 3 + (tw == 16 ? 0 : (tw == 32 ? 1 : 2));
    if (!c->grid[k]) c->grid[k] = halfstep_grid(maxs, tw);
    return c->grid[k];
}

// The half-step kernel for more than 16 sources keeps each workgroup's
// view-selection costs (32 sources x 256 threads) in global memory; sized for
// the largest grid it is launched with.  Called before any launch (and before
// graph capture).
int ensure_c0(vsm_ctx *c, int maxs, int tw)
{
    if (maxs <= 16) return VSM_OK;
    const size_t need = (size_t)grid_for(c, maxs, tw) * 32 * 256;
    if (c->c0_floats >= need) return VSM_OK;
    if (c->d_c0) dfree(c, c->d_c0);
    c->d_c0 = nullptr;
    c->c0_floats = 0;
    const int rc = dalloc(c, &c->d_c0, need);
    if (rc) return rc;
    c->c0_floats = need;
    return VSM_OK;
}

int record_halfstep(vsm_ctx *c, int maxs, int tw, const ViewDesc *dviews, const Tile *dtiles, int ntiles,
                    const PMConst &pc, bool timed)
{
    {
        const int rc = ensure_c0(c, maxs, tw);
        if (rc) return rc;
    }
    if (timed) HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, launch_halfstep(maxs, tw, dviews, dtiles, ntiles, pc, c->d_counters, c->d_work,
                              grid_for(c, maxs, tw), c->d_c0, c->stream));
    if (timed) {
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        HIPCHK(c, hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->halfstep_ms += ms;
    }
    c->halfstep_launches += 1;
    return VSM_OK;
}

int read_counters(vsm_ctx *c, const PMConst &pc)
{
    Counters h;
    HIPCHK(c, hipMemcpyAsync(&h, c->d_counters, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, launch_fill_u64(reinterpret_cast<unsigned long long *>(c->d_counters), 4, 0ull, c->stream));
    c->acc_ncc += (int64_t)h.ncc_evals;
    c->acc_geo += (int64_t)h.geo_evals;
    c->acc_taps += (int64_t)h.ncc_evals * pc.ntaps * pc.ntaps;
    return VSM_OK;
}

int finalize_outputs(vsm_ctx *c, ViewState &vs, int pass, int ref_id, vsm_view_out *out,
                     const uint32_t *d_mask)
{
    ImageRec &ref = c->images[ref_id];
    const size_t hw = (size_t)ref.W * ref.H;
    ViewDesc *dv = nullptr;
    float *tmp = nullptr;
    int rc = dalloc(c, &dv, 1);
    if (!rc) rc = dalloc(c, &tmp, hw * 2);
    if (rc) { dfree(c, dv); return rc; }
    if ((!ref.depth[pass] && (rc = dalloc(c, &ref.depth[pass], hw))) ||
        (!ref.normal[pass] && (rc = dalloc(c, &ref.normal[pass], hw * 3)))) {
        dfree(c, dv);
        dfree(c, tmp);
        return rc;
    }
    HIPCHK(c, hipMemcpyAsync(dv, &vs.desc, sizeof(ViewDesc), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_finalize(dv, 0, (int)hw, tmp, ref.normal[pass], tmp + hw, ref.depth[pass],
                              c->stream));
    if (pass == 0) ++ref.depth_ver;
    if (out && out->depth)
        HIPCHK(c, hipMemcpyAsync(out->depth, tmp, hw * 4, hipMemcpyDeviceToHost, c->stream));
    if (out && out->normal)
        HIPCHK(c, hipMemcpyAsync(out->normal, ref.normal[pass], hw * 12, hipMemcpyDeviceToHost,
                                 c->stream));
    if (out && out->cost)
        HIPCHK(c, hipMemcpyAsync(out->cost, tmp + hw, hw * 4, hipMemcpyDeviceToHost, c->stream));
    if (out && out->consistency) {
        if (d_mask) HIPCHK(c, hipMemcpyAsync(out->consistency, d_mask, hw * 4, hipMemcpyDeviceToHost, c->stream));
        else memset(out->consistency, 0, hw * 4);
    }
    // dv and tmp go back to the cache; the copies above are ordered before
    // any later use of them on the stream, and the caller synchronises
    dfree(c, dv);
    dfree(c, tmp);
    return VSM_OK;
}

// ---------------------------------------------------------------------------
// Staged batched solve.  A batch holds n prepared views; each step runs
// half-steps over all of them with one persistent launch per half-step
// (optionally replayed from a captured hipGraph, kept for later steps of the
// same length); the end runs the filter and writes the outputs.  vsm_solve is
// begin + every half-step + end.
// ---------------------------------------------------------------------------
void batch_release(vsm_ctx *c, Batch &b)
{
    (void)hipStreamSynchronize(c->stream);
    for (auto &g : b.graphs) (void)hipGraphExecDestroy(g.second);
    b.graphs.clear();
    for (auto &v : b.views) release_view(c, v);
    b.views.clear();
    dfree(c, b.dviews);
    dfree(c, b.dtiles);
    b.dviews = nullptr;
    b.dtiles = nullptr;
}

// Prepare the views, upload their descriptors and the concatenated tile list,
// and set the initial hypotheses: random for the photometric pass; for the
// geometric pass the depth/normal maps the request supplies, else the
// image's photometric depth and normal maps on the device (kept by its
// photometric solve or set from files).
int batch_begin(vsm_ctx *c, const vsm_view_req *reqs, int n, int pass, int use_graph, Batch &b)
{
    b.pass = pass;
    b.use_graph = use_graph != 0;
    b.params = c->params;
    b.views.resize(n);
    int rc = VSM_OK, prepared = 0;
    for (int v = 0; v < n && !rc; ++v) {
        rc = prepare_view(c, reqs[v], pass, (uint32_t)reqs[v].ref_id, b.views[v]);
        if (!rc) ++prepared;
    }
    if (rc) {
        b.views.resize(prepared);
        batch_release(c, b);
        return rc;
    }
    std::vector<Tile> tiles;
    std::vector<ViewDesc> hv(n);
    for (int v = 0; v < n; ++v) {
        b.maxs = std::max(b.maxs, b.views[v].desc.n_src);
        tiles_for(b.views[v].desc, v, tiles, b.params.tile_width);
        b.ref_ids.push_back(reqs[v].ref_id);
        hv[v] = b.views[v].desc;
    }
    b.ntiles = (int)tiles.size();
    b.nsteps = 2 * b.params.num_iterations;
    if (!tile_width_supported(b.params.tile_width, b.maxs)) {
        batch_release(c, b);
        return fail(c, VSM_EINVAL, "tile_width %d is not available for %d sources (64-wide tiles take at "
                    "most 16)", b.params.tile_width, b.maxs);
    }
    rc = dalloc(c, &b.dviews, n);
    if (!rc) rc = dalloc(c, &b.dtiles, tiles.size());
    if (!rc) rc = ensure_c0(c, b.maxs, b.params.tile_width);     // before any capture
    if (rc) {
        batch_release(c, b);
        return rc;
    }
#define BEGINCHK(expr)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            batch_release(c, b);                                                                \
            return fail(c, VSM_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));        \
        }                                                                                       \
    } while (0)
    const PMConst pc = make_const(b.params);
    BEGINCHK(hipMemcpyAsync(b.dviews, hv.data(), n * sizeof(ViewDesc), hipMemcpyHostToDevice, c->stream));
    BEGINCHK(hipMemcpyAsync(b.dtiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice,
                            c->stream));
    if (pass == 0) {
        BEGINCHK(launch_init(b.dviews, b.dtiles, b.ntiles, pc, 1, b.params.tile_width, c->stream));
    } else {
        for (int v = 0; v < n; ++v) {
            const vsm_view_req &r = reqs[v];
            const ImageRec &ref = c->images[r.ref_id];
            const size_t hw = (size_t)ref.W * ref.H;
            if (r.init_depth && r.init_normal) {
                float *tmp = nullptr;
                if ((rc = dalloc(c, &tmp, hw * 4))) {
                    batch_release(c, b);
                    return rc;
                }
                BEGINCHK(hipMemcpyAsync(tmp, r.init_depth, hw * 4, hipMemcpyHostToDevice, c->stream));
                BEGINCHK(hipMemcpyAsync(tmp + hw, r.init_normal, hw * 12, hipMemcpyHostToDevice, c->stream));
                BEGINCHK(launch_load_hyp(b.dviews, v, (int)hw, tmp, tmp + hw, pc.cos_max_normal, c->stream));
                BEGINCHK(hipStreamSynchronize(c->stream));
                dfree(c, tmp);
            } else if (ref.depth[0] && ref.normal[0]) {
                BEGINCHK(launch_load_hyp(b.dviews, v, (int)hw, ref.depth[0], ref.normal[0],
                                         pc.cos_max_normal, c->stream));
            } else {
                batch_release(c, b);
                return fail(c, VSM_EINVAL,
                            "geometric pass of image %d needs its photometric depth and normal maps "
                            "(solve pass 0 first, set them, or pass init_depth/init_normal)", r.ref_id);
            }
        }
        BEGINCHK(launch_init(b.dviews, b.dtiles, b.ntiles, pc, 0, b.params.tile_width, c->stream));
    }
    BEGINCHK(hipStreamSynchronize(c->stream));
#undef BEGINCHK
    return VSM_OK;
}

// Run up to k more half-steps (fewer when the batch has fewer left).
int batch_step(vsm_ctx *c, Batch &b, int k)
{
    const int m = std::min(k, b.nsteps - b.steps_done);
    if (m <= 0) return VSM_OK;
    const PMConst pc = make_const(b.params);
    if (b.use_graph) {
        hipGraphExec_t exec = nullptr;
        for (auto &g : b.graphs)
            if (g.first == m) exec = g.second;
        if (!exec) {
            // [zero the tile cursors, half-step, advance the device step
            // counters] x m, captured once per length and replayed
            hipGraph_t graph = nullptr;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t ce = hipSuccess;
            for (int s = 0; s < m && ce == hipSuccess; ++s) {
                ce = launch_halfstep(b.maxs, b.params.tile_width, b.dviews, b.dtiles, b.ntiles, pc,
                                     c->d_counters, c->d_work, grid_for(c, b.maxs, b.params.tile_width),
                                     c->d_c0, c->stream);
                if (ce == hipSuccess) ce = launch_advance(b.dviews, (int)b.views.size(), c->stream);
            }
            const hipError_t ee = hipStreamEndCapture(c->stream, &graph);
            if (ce != hipSuccess || ee != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return fail(c, VSM_EDEVICE, "graph capture failed: %s",
                            hipGetErrorString(ce != hipSuccess ? ce : ee));
            }
            const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess)
                return fail(c, VSM_EDEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(ie));
            b.graphs.emplace_back(m, exec);
        }
        if (c->timing) HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        HIPCHK(c, hipGraphLaunch(exec, c->stream));
        if (c->timing) {
            HIPCHK(c, hipEventRecord(c->ev1, c->stream));
            HIPCHK(c, hipEventSynchronize(c->ev1));
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
            c->halfstep_ms += ms;
        }
        c->halfstep_launches += m;
    } else {
        for (int s = 0; s < m; ++s) {
            const int rc = record_halfstep(c, b.maxs, b.params.tile_width, b.dviews, b.dtiles, b.ntiles, pc,
                                           c->timing);
            if (rc) return rc;
            HIPCHK(c, launch_advance(b.dviews, (int)b.views.size(), c->stream));
        }
    }
    b.steps_done += m;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return read_counters(c, pc);
}

// Filter (when the batch's parameters ask for it) and outputs of every view.
int batch_finish(vsm_ctx *c, Batch &b, vsm_view_out *outs)
{
    const int n = (int)b.views.size();
    const PMConst pc = make_const(b.params);
    uint32_t *d_masks = nullptr;
    std::vector<size_t> mask_off(n, 0);
    int rc = VSM_OK;
    if (b.params.filter) {
        size_t tot = 0;
        for (int v = 0; v < n; ++v) {
            mask_off[v] = tot;
            tot += (size_t)b.views[v].desc.W * b.views[v].desc.H;
        }
        if ((rc = dalloc(c, &d_masks, tot))) return rc;
        // host tile lists stay alive until the stream is synchronised below
        std::vector<std::vector<Tile>> filter_tiles(n);
        for (int v = 0; v < n && !rc; ++v) {
            // one launch per view so the mask offsets stay simple
            std::vector<Tile> &vt = filter_tiles[v];
            tiles_for(b.views[v].desc, 0, vt);
            Tile *dt = nullptr;
            if ((rc = dalloc(c, &dt, vt.size()))) break;
            hipError_t e = hipMemcpyAsync(dt, vt.data(), vt.size() * sizeof(Tile), hipMemcpyHostToDevice,
                                          c->stream);
            if (e == hipSuccess)
                e = launch_filter(b.maxs, b.dviews + v, dt, (int)vt.size(), pc, d_masks + mask_off[v],
                                  c->stream);
            dfree(c, dt);                      // back to the cache, stream-ordered
            if (e != hipSuccess) rc = fail(c, VSM_EDEVICE, "filter launch failed: %s", hipGetErrorString(e));
        }
        const hipError_t se = hipStreamSynchronize(c->stream);
        if (!rc && se != hipSuccess)
            rc = fail(c, VSM_EDEVICE, "hipStreamSynchronize failed: %s", hipGetErrorString(se));
    }
    for (int v = 0; v < n && !rc; ++v)
        rc = finalize_outputs(c, b.views[v], b.pass, b.ref_ids[v], outs ? &outs[v] : nullptr,
                              d_masks ? d_masks + mask_off[v] : nullptr);
    const hipError_t se = hipStreamSynchronize(c->stream);
    if (!rc && se != hipSuccess)
        rc = fail(c, VSM_EDEVICE, "hipStreamSynchronize failed: %s", hipGetErrorString(se));
    dfree(c, d_masks);
    return rc;
}

}  // namespace

// ===========================================================================
// Public C ABI
// ===========================================================================
extern "C" {

void vsm_default_params(vsm_params *p)
{
    if (!p) return;
    memset(p, 0, sizeof *p);
    p->window_radius = 3;             // stereo_pipeline.py:444
    p->window_step = 1;
    p->sigma_spatial = -1.f;
    p->sigma_color = 0.2f;
    p->ncc_sigma = 0.6f;
    p->num_samples = 15;              // colmap_mvs_commands.py:46
    p->num_iterations = 12;           // colmap_mvs_commands.py:47
    p->min_triangulation_angle = 5.f; // colmap_mvs_commands.py:42
    p->incident_angle_sigma = 0.9f;
    p->geom_consistency = 0;          // colmap_mvs_commands.py:44
    p->geom_consistency_regularizer = 0.3f;    // colmap_mvs_commands.py:66
    p->geom_consistency_max_cost = 3.f;        // colmap_mvs_commands.py:67
    p->filter = 0;                             // colmap_mvs_commands.py:43
    p->filter_min_ncc = 0.1f;                  // colmap_mvs_commands.py:70
    p->filter_min_triangulation_angle = 4.99f; // colmap_mvs_commands.py:69
    p->filter_geom_consistency_max_cost = 1.f; // colmap_mvs_commands.py:71
    p->filter_min_num_consistent = 2;          // colmap_mvs_commands.py:72
    p->depth_min = -1.f;
    p->depth_max = -1.f;
    p->perturbation = 0.05f;
    p->normal_perturbation = 0.3f;
    p->max_normal_angle = 80.f;
    p->perturb_decay = 0.5f;
    p->n_far = 6;
    p->seed = 0x5A7E11173Dull;
    p->draw_group = 0;
    p->tile_width = 32;
}

int vsm_create(int device, uint64_t hbm_budget_bytes, vsm_ctx **out)
{
    if (!out) return VSM_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return VSM_EDEVICE;
    if (device < 0 || device >= ndev) return VSM_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return VSM_EDEVICE;
    // device address of the allocation-free scratch, once per device
    if (scratch_init(device) != hipSuccess) return VSM_EDEVICE;
    auto *c = new vsm_ctx();
    c->device = device;
    c->budget = hbm_budget_bytes;
    vsm_default_params(&c->params);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return VSM_EDEVICE;
    }
    if (dalloc(c, &c->d_counters, 1) != VSM_OK || dalloc(c, &c->d_work, 8) != VSM_OK) {
        delete c;
        return VSM_ENOMEM;
    }
    (void)launch_fill_u64(reinterpret_cast<unsigned long long *>(c->d_counters), 4, 0ull, c->stream);
    (void)hipStreamSynchronize(c->stream);
    *out = c;
    return VSM_OK;
}

int vsm_destroy(vsm_ctx *c)
{
    if (!c) return VSM_EINVAL;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &kv : c->batches)
        for (auto &g : kv.second->graphs) (void)hipGraphExecDestroy(g.second);
    c->batches.clear();
    for (auto &kv : c->allocs) (void)hipFree(kv.first);
    c->allocs.clear();
    for (auto &kv : c->cache) (void)hipFree(kv.second);
    c->cache.clear();
    c->cached_bytes = 0;
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return VSM_OK;
}

const char *vsm_last_error(vsm_ctx *c) { return c ? c->err.c_str() : "null context"; }

int vsm_set_params(vsm_ctx *c, const vsm_params *p)
{
    if (!c || !p) return VSM_EINVAL;
    int rc = check_params(c, *p);
    if (rc) return rc;
    c->params = *p;
    return VSM_OK;
}

int vsm_get_params(vsm_ctx *c, vsm_params *p)
{
    if (!c || !p) return VSM_EINVAL;
    *p = c->params;
    return VSM_OK;
}

int vsm_add_image(vsm_ctx *c, int32_t id, const uint8_t *gray, int32_t w, int32_t h, int32_t stride,
                  const double P4[16])
{
    if (!c) return VSM_EINVAL;
    if (!P4) return fail(c, VSM_EINVAL, "NULL P4");
    if (w < 2 || h < 2 || w > 32768 || h > 32768 || (int64_t)w * h > (1ll << 30))
        return fail(c, VSM_EINVAL, "image size %dx%d unsupported", w, h);
    if (stride < w) return fail(c, VSM_EINVAL, "row_stride %d < width %d", stride, w);
    for (int k = 0; k < 16; ++k)
        if (!std::isfinite(P4[k])) return fail(c, VSM_EINVAL, "P4 has non-finite entries");
    HIPCHK(c, hipSetDevice(c->device));
    if (c->images.count(id)) {               // replace: the old image must not be in use
        const int rc = vsm_remove_image(c, id);
        if (rc) return rc;
    }
    ImageRec r;
    r.id = id;
    r.W = w;
    r.H = h;
    memcpy(r.P4, P4, sizeof r.P4);
    if (!inv4(P4, r.Pinv)) return fail(c, VSM_EINVAL, "P4 of image %d is singular", id);
    if (gray) {
        int rc = dalloc(c, &r.gray, (size_t)w * h);
        if (!rc) rc = dalloc(c, &r.pair, (size_t)(w + 1) * h);
        if (rc) { dfree(c, r.gray); return rc; }
        HIPCHK(c, hipMemcpy2DAsync(r.gray, w, gray, stride, w, h, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, launch_pack_pair(r.gray, w, h, r.pair, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    c->images[id] = r;
    return VSM_OK;
}

int vsm_remove_image(vsm_ctx *c, int32_t id)
{
    if (!c) return VSM_EINVAL;
    auto it = c->images.find(id);
    if (it == c->images.end()) return VSM_ENOENT;
    {
        std::lock_guard<std::mutex> lk(c->serve_mu);
        auto uses = [id](const ServeReq &q) {
            return q.ref == id || std::find(q.srcs.begin(), q.srcs.end(), id) != q.srcs.end();
        };
        for (const ServeReq *q : c->active)
            if (uses(*q)) return fail(c, VSM_EINVAL, "image %d is used by an active served request", id);
        for (const auto &q : c->queue)
            if (uses(*q)) return fail(c, VSM_EINVAL, "image %d is used by a queued served request", id);
    }
    for (const auto &kv : c->batches)
        for (int rid : kv.second->ref_ids)
            if (rid == id) return fail(c, VSM_EINVAL, "image %d is a reference of open batch %d", id, kv.first);
    (void)hipStreamSynchronize(c->stream);
    ImageRec &r = it->second;
    dfree(c, r.gray);
    dfree(c, r.pair);
    dfree(c, r.depth[0]);
    dfree(c, r.depth[1]);
    dfree(c, r.normal[0]);
    dfree(c, r.normal[1]);
    dfree(c, r.depth4);
    c->images.erase(it);
    return VSM_OK;
}

int vsm_set_depth_map(vsm_ctx *c, int32_t id, int32_t kind, const float *depth, int32_t w, int32_t h)
{
    if (!c || !depth) return VSM_EINVAL;
    if (kind != 0 && kind != 1) return fail(c, VSM_EINVAL, "depth map kind must be 0 or 1");
    auto it = c->images.find(id);
    if (it == c->images.end()) return fail(c, VSM_ENOENT, "unknown image %d", id);
    ImageRec &r = it->second;
    if (w != r.W || h != r.H) return fail(c, VSM_EINVAL, "depth map %dx%d != image %dx%d", w, h, r.W, r.H);
    if (!r.depth[kind]) {
        int rc = dalloc(c, &r.depth[kind], (size_t)w * h);
        if (rc) return rc;
    }
    HIPCHK(c, hipMemcpyAsync(r.depth[kind], depth, (size_t)w * h * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (kind == 0) ++r.depth_ver;
    return VSM_OK;
}

int vsm_set_normal_map(vsm_ctx *c, int32_t id, int32_t kind, const float *normal, int32_t w,
                       int32_t h)
{
    if (!c || !normal) return VSM_EINVAL;
    if (kind != 0 && kind != 1) return fail(c, VSM_EINVAL, "normal map kind must be 0 or 1");
    auto it = c->images.find(id);
    if (it == c->images.end()) return fail(c, VSM_ENOENT, "unknown image %d", id);
    ImageRec &r = it->second;
    if (w != r.W || h != r.H) return fail(c, VSM_EINVAL, "normal map %dx%d != image %dx%d", w, h, r.W, r.H);
    HIPCHK(c, hipSetDevice(c->device));
    if (!r.normal[kind]) {
        int rc = dalloc(c, &r.normal[kind], (size_t)w * h * 3);
        if (rc) return rc;
    }
    HIPCHK(c, hipMemcpyAsync(r.normal[kind], normal, (size_t)w * h * 12, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VSM_OK;
}

int vsm_get_depth_map(vsm_ctx *c, int32_t id, int32_t kind, float *depth, int32_t w, int32_t h)
{
    if (!c || !depth) return VSM_EINVAL;
    if (kind != 0 && kind != 1) return fail(c, VSM_EINVAL, "depth map kind must be 0 or 1");
    auto it = c->images.find(id);
    if (it == c->images.end()) return fail(c, VSM_ENOENT, "unknown image %d", id);
    const ImageRec &r = it->second;
    if (w != r.W || h != r.H) return fail(c, VSM_EINVAL, "buffer %dx%d != image %dx%d", w, h, r.W, r.H);
    if (!r.depth[kind]) return fail(c, VSM_ENOENT, "image %d has no depth map of kind %d", id, kind);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(depth, r.depth[kind], (size_t)w * h * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VSM_OK;
}

int vsm_solve(vsm_ctx *c, const vsm_view_req *reqs, int32_t n, int32_t pass, vsm_view_out *outs,
              int32_t use_graph)
{
    if (!c) return VSM_EINVAL;
    if (!reqs || n <= 0) return fail(c, VSM_EINVAL, "no requests");
    if (pass != 0 && pass != 1) return fail(c, VSM_EINVAL, "pass must be 0 or 1");
    HIPCHK(c, hipSetDevice(c->device));
    Batch b;
    int rc = batch_begin(c, reqs, n, pass, use_graph, b);
    if (rc) return rc;
    rc = batch_step(c, b, b.nsteps);
    if (!rc) rc = batch_finish(c, b, outs);
    batch_release(c, b);
    return rc;
}

int vsm_batch_begin(vsm_ctx *c, const vsm_view_req *reqs, int32_t n, int32_t pass, int32_t use_graph,
                    int32_t *batch)
{
    if (!c || !batch) return VSM_EINVAL;
    *batch = 0;
    if (!reqs || n <= 0) return fail(c, VSM_EINVAL, "no requests");
    if (pass != 0 && pass != 1) return fail(c, VSM_EINVAL, "pass must be 0 or 1");
    HIPCHK(c, hipSetDevice(c->device));
    auto b = std::make_unique<Batch>();
    const int rc = batch_begin(c, reqs, n, pass, use_graph, *b);
    if (rc) return rc;
    const int id = c->next_batch++;
    c->batches[id] = std::move(b);
    *batch = id;
    return VSM_OK;
}

int vsm_batch_step(vsm_ctx *c, int32_t batch, int32_t n_halfsteps, int32_t *remaining)
{
    if (!c) return VSM_EINVAL;
    auto it = c->batches.find(batch);
    if (it == c->batches.end()) return fail(c, VSM_ENOENT, "unknown batch %d", batch);
    if (n_halfsteps < 0) return fail(c, VSM_EINVAL, "negative half-step count");
    HIPCHK(c, hipSetDevice(c->device));
    Batch &b = *it->second;
    const int rc = batch_step(c, b, n_halfsteps);
    if (remaining) *remaining = b.nsteps - b.steps_done;
    return rc;
}

int vsm_batch_end(vsm_ctx *c, int32_t batch, vsm_view_out *outs, int32_t discard)
{
    if (!c) return VSM_EINVAL;
    auto it = c->batches.find(batch);
    if (it == c->batches.end()) return fail(c, VSM_ENOENT, "unknown batch %d", batch);
    Batch &b = *it->second;
    if (!discard && b.steps_done < b.nsteps)
        return fail(c, VSM_EINVAL, "batch %d has %d half-steps left", batch, b.nsteps - b.steps_done);
    HIPCHK(c, hipSetDevice(c->device));
    const int rc = discard ? VSM_OK : batch_finish(c, b, outs);
    batch_release(c, b);
    c->batches.erase(it);
    return rc;
}

int vsm_eval_cost(vsm_ctx *c, const vsm_view_req *req, int32_t pass, const float *depth,
                  const float *normal, float *out)
{
    if (!c || !req || !depth || !normal || !out) return VSM_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    ViewState vs;
    int rc = prepare_view(c, *req, pass, (uint32_t)req->ref_id, vs);
    if (rc) return rc;
    const ViewDesc &V = vs.desc;
    const size_t hw = (size_t)V.W * V.H;
    std::vector<float4> h(hw);
    for (size_t i = 0; i < hw; ++i)
        h[i] = make_float4(depth[i], normal[3 * i], normal[3 * i + 1], normal[3 * i + 2]);
    std::vector<Tile> tiles;
    tiles_for(V, 0, tiles);
    ViewDesc *dv = nullptr;
    Tile *dt = nullptr;
    float *dout = nullptr;
    rc = dalloc(c, &dv, 1);
    if (!rc) rc = dalloc(c, &dt, tiles.size());
    if (!rc) rc = dalloc(c, &dout, hw * V.n_src);
    if (!rc) {
        HIPCHK(c, hipMemcpyAsync(vs.hyp, h.data(), hw * sizeof(float4), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(dv, &V, sizeof V, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(dt, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, launch_eval_cost(V.n_src, dv, dt, (int)tiles.size(), make_const(c->params), dout, c->stream));
        HIPCHK(c, hipMemcpyAsync(out, dout, hw * V.n_src * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    dfree(c, dv); dfree(c, dt); dfree(c, dout);
    release_view(c, vs);
    return rc;
}

int vsm_eval_cost_staging(vsm_ctx *c, const vsm_view_req *req, int32_t pass, const float *depth,
                          const float *normal, int32_t stage, int32_t repeats, float *out,
                          float *ms_per_launch, float *staged_fraction)
{
    if (!c || !req || !depth || !normal || repeats < 1) return VSM_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    ViewState vs;
    int rc = prepare_view(c, *req, pass, (uint32_t)req->ref_id, vs);
    if (rc) return rc;
    const ViewDesc &V = vs.desc;
    const size_t hw = (size_t)V.W * V.H;
    std::vector<float4> h(hw);
    for (size_t i = 0; i < hw; ++i)
        h[i] = make_float4(depth[i], normal[3 * i], normal[3 * i + 1], normal[3 * i + 2]);
    std::vector<Tile> tiles;
    tiles_for(V, 0, tiles);
    ViewDesc *dv = nullptr;
    Tile *dt = nullptr;
    float *dout = nullptr;
    unsigned int *dn = nullptr;
    rc = dalloc(c, &dv, 1);
    if (!rc) rc = dalloc(c, &dt, tiles.size());
    if (!rc) rc = dalloc(c, &dout, hw * V.n_src);
    if (!rc) rc = dalloc(c, &dn, 1);
    if (!rc) {
        const PMConst pc = make_const(c->params);
        const int nt = (int)tiles.size();
        HIPCHK(c, hipMemcpyAsync(vs.hyp, h.data(), hw * sizeof(float4), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(dv, &V, sizeof V, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(dt, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, launch_eval_stage(V.n_src, stage, dv, dt, nt, pc, dout, nullptr, c->stream));   // warm
        HIPCHK(c, launch_fill_u32(dn, 1, 0u, c->stream));
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        for (int r = 0; r < repeats; ++r)
            HIPCHK(c, launch_eval_stage(V.n_src, stage, dv, dt, nt, pc, dout, r == 0 ? dn : nullptr, c->stream));
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        HIPCHK(c, hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        unsigned int ns = 0;
        HIPCHK(c, hipMemcpyAsync(out, dout, hw * V.n_src * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(&ns, dn, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (ms_per_launch) *ms_per_launch = ms / repeats;
        if (staged_fraction) *staged_fraction = (float)ns / (float)((size_t)nt * V.n_src);
    }
    dfree(c, dv); dfree(c, dt); dfree(c, dout); dfree(c, dn);
    release_view(c, vs);
    return rc;
}

static int debug_state(vsm_ctx *c, const vsm_view_req *req, int pass, int mode, int iter, int color,
                       float *depth, float *normal, float *cost, float *prob, uint32_t *mask)
{
    if (!c || !req || !depth || !normal) return VSM_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    ViewState vs;
    int rc = prepare_view(c, *req, pass, (uint32_t)req->ref_id, vs);
    if (rc) return rc;
    ViewDesc &V = vs.desc;
    const size_t hw = (size_t)V.W * V.H;
    V.step = 2 * iter + color;
    V.num_steps = V.step + 1;
    const int tw = c->params.tile_width;
    if (mode == 1 && !tile_width_supported(tw, V.n_src)) {
        release_view(c, vs);
        return fail(c, VSM_EINVAL, "tile_width %d is not available for %d sources", tw, V.n_src);
    }
    std::vector<Tile> tiles;
    tiles_for(V, 0, tiles, mode == 2 ? kTileW : tw);
    ViewDesc *dv = nullptr;
    Tile *dt = nullptr;
    uint32_t *dm = nullptr;
    rc = dalloc(c, &dv, 1);
    if (!rc) rc = dalloc(c, &dt, tiles.size());
    if (!rc && mode == 2) rc = dalloc(c, &dm, hw);
    if (!rc && mode == 1) rc = ensure_c0(c, V.n_src, tw);
    const PMConst pc = make_const(c->params);
    std::vector<float4> h(hw);
    if (!rc) {
        if (mode != 0) {
            for (size_t i = 0; i < hw; ++i)
                h[i] = make_float4(depth[i], normal[3 * i], normal[3 * i + 1], normal[3 * i + 2]);
            HIPCHK(c, hipMemcpyAsync(vs.hyp, h.data(), hw * sizeof(float4), hipMemcpyHostToDevice, c->stream));
            if (cost) HIPCHK(c, hipMemcpyAsync(vs.cost, cost, hw * 4, hipMemcpyHostToDevice, c->stream));
            if (prob) HIPCHK(c, hipMemcpyAsync(vs.prob, prob, hw * V.n_src * 4, hipMemcpyHostToDevice, c->stream));
        }
        HIPCHK(c, hipMemcpyAsync(dv, &V, sizeof V, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(dt, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, c->stream));
        const int nt = (int)tiles.size();
        if (mode == 0) HIPCHK(c, launch_init(dv, dt, nt, pc, 1, tw, c->stream));
        else if (mode == 1)
            HIPCHK(c, launch_halfstep(V.n_src, tw, dv, dt, nt, pc, c->d_counters, c->d_work,
                                      grid_for(c, V.n_src, tw), c->d_c0, c->stream));
        else HIPCHK(c, launch_filter(V.n_src, dv, dt, nt, pc, dm, c->stream));
        HIPCHK(c, hipMemcpyAsync(h.data(), vs.hyp, hw * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
        if (cost && mode != 2) HIPCHK(c, hipMemcpyAsync(cost, vs.cost, hw * 4, hipMemcpyDeviceToHost, c->stream));
        if (prob && mode != 2) HIPCHK(c, hipMemcpyAsync(prob, vs.prob, hw * V.n_src * 4, hipMemcpyDeviceToHost, c->stream));
        if (mask && mode == 2) HIPCHK(c, hipMemcpyAsync(mask, dm, hw * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < hw; ++i) {
            depth[i] = h[i].x;
            normal[3 * i] = h[i].y; normal[3 * i + 1] = h[i].z; normal[3 * i + 2] = h[i].w;
        }
        if (mode == 1) rc = read_counters(c, pc);
    }
    dfree(c, dv); dfree(c, dt); dfree(c, dm);
    release_view(c, vs);
    return rc;
}

int vsm_debug_init(vsm_ctx *c, const vsm_view_req *req, int32_t pass, float *depth, float *normal,
                   float *cost, float *prob)
{
    return debug_state(c, req, pass, 0, 0, 0, depth, normal, cost, prob, nullptr);
}

int vsm_debug_halfstep(vsm_ctx *c, const vsm_view_req *req, int32_t pass, int32_t iter, int32_t color,
                       float *depth, float *normal, float *cost, float *prob)
{
    if (!cost || !prob) return VSM_EINVAL;
    if (iter < 0 || color < 0 || color > 1) return VSM_EINVAL;
    return debug_state(c, req, pass, 1, iter, color, depth, normal, cost, prob, nullptr);
}

int vsm_debug_filter(vsm_ctx *c, const vsm_view_req *req, int32_t pass, float *depth, float *normal,
                     uint32_t *mask)
{
    return debug_state(c, req, pass, 2, 0, 0, depth, normal, nullptr, nullptr, mask);
}

// ---------------------------------------------------------------------------
// Fusion
// ---------------------------------------------------------------------------
static DsmGridDev grid_dev(const vsm_dsm_grid &g)
{
    DsmGridDev d;
    d.ul_e = g.ul_e; d.ul_n = g.ul_n; d.e_res = g.e_res; d.n_res = g.n_res;
    d.e_size = g.e_size; d.n_size = g.n_size;
    return d;
}

static int check_grid(vsm_ctx *c, const vsm_dsm_grid *g)
{
    if (!g) return fail(c, VSM_EINVAL, "NULL grid");
    if (g->e_size <= 0 || g->n_size <= 0 || (int64_t)g->e_size * g->n_size > (1ll << 31))
        return fail(c, VSM_EINVAL, "grid size %dx%d unsupported", g->e_size, g->n_size);
    if (!(g->e_res > 0) || !(g->n_res > 0)) return fail(c, VSM_EINVAL, "grid resolution must be > 0");
    return VSM_OK;
}

int vsm_fuse_dsm(vsm_ctx *c, const vsm_dsm_grid *grid, const int32_t *ids, int32_t n,
                 int32_t depth_kind, int32_t min_measurements, const double *inv_proj_mats,
                 float *dsm_out, float *per_view_out)
{
    if (!c) return VSM_EINVAL;
    int rc = check_grid(c, grid);
    if (rc) return rc;
    if (!ids || n <= 0 || n > 64) return fail(c, VSM_EINVAL, "1..64 images required, got %d", n);
    if (depth_kind != 0 && depth_kind != 1) return fail(c, VSM_EINVAL, "depth_kind must be 0 or 1");
    if (!dsm_out) return fail(c, VSM_EINVAL, "NULL dsm_out");
    for (int v = 0; v < n; ++v) {
        auto it = c->images.find(ids[v]);
        if (it == c->images.end()) return fail(c, VSM_ENOENT, "unknown image %d", ids[v]);
        if (!it->second.depth[depth_kind])
            return fail(c, VSM_EINVAL, "image %d has no %s depth map", ids[v],
                        depth_kind ? "geometric" : "photometric");
    }
    HIPCHK(c, hipSetDevice(c->device));
    const DsmGridDev g = grid_dev(*grid);
    const long long ncell = (long long)g.e_size * g.n_size;
    unsigned long long *cells = nullptr;
    float *stack = nullptr, *tmp = nullptr, *fused = nullptr;
    double *mats = nullptr;
    rc = dalloc(c, &cells, 2 * ncell);        // two key grids, used by alternate views
    if (!rc) rc = dalloc(c, &stack, ncell * n);
    if (!rc) rc = dalloc(c, &tmp, ncell);
    if (!rc) rc = dalloc(c, &fused, ncell);
    if (!rc) rc = dalloc(c, &mats, 16 * (size_t)n);
    if (rc) { dfree(c, cells); dfree(c, stack); dfree(c, tmp); dfree(c, fused); return rc; }
    std::vector<double> hm(16 * (size_t)n);
    for (int v = 0; v < n; ++v)
        memcpy(&hm[16 * v], inv_proj_mats ? inv_proj_mats + 16 * v : c->images[ids[v]].Pinv,
               16 * sizeof(double));
    HIPCHK(c, hipMemcpyAsync(mats, hm.data(), hm.size() * 8, hipMemcpyHostToDevice, c->stream));
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    int64_t pixels = 0;
    // view v rasterises into key grid v % 2; its per-view DSM pass zeroes
    // the other grid for view v + 1
    HIPCHK(c, launch_fill_u64n(cells, ncell, 0ull, c->stream));
    for (int v = 0; v < n; ++v) {
        const ImageRec &r = c->images[ids[v]];
        unsigned long long *kv = cells + (size_t)(v & 1) * ncell, *knext = cells + (size_t)((v + 1) & 1) * ncell;
        HIPCHK(c, launch_rasterize_view(r.depth[depth_kind], r.W, r.H, mats + 16 * v, g,
                                        c->has_geo ? &c->geo : nullptr, kv, c->stream));
        HIPCHK(c, launch_view_dsm(kv, knext, g.e_size, g.n_size, stack + (size_t)v * ncell, c->stream));
        pixels += (int64_t)r.W * r.H;
    }
    HIPCHK(c, launch_fuse_stack(stack, n, ncell, min_measurements, tmp, c->stream));
    HIPCHK(c, launch_median3(tmp, g.e_size, g.n_size, fused, c->stream));
    if (c->timing) {
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        HIPCHK(c, hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->fuse_ms += ms;
    }
    c->fuse_launches += 1;
    c->fuse_pixels += pixels;
    HIPCHK(c, hipMemcpyAsync(dsm_out, fused, ncell * 4, hipMemcpyDeviceToHost, c->stream));
    if (per_view_out)
        HIPCHK(c, hipMemcpyAsync(per_view_out, stack, ncell * 4 * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c, cells); dfree(c, stack); dfree(c, tmp); dfree(c, fused); dfree(c, mats);
    return VSM_OK;
}

void vsm_default_fuse3d_params(vsm_fuse3d_params *p)
{
    if (!p) return;
    p->min_num_pixels = 4;           // aggregate_3d.py:47-50
    p->max_reproj_error = 2.f;
    p->max_depth_error = 1.f;
    p->max_normal_error_deg = 10.f;
}

int vsm_fuse_points(vsm_ctx *c, const int32_t *ids, int32_t n, int32_t depth_kind,
                    const vsm_fuse3d_params *params, const double *inv_proj_mats,
                    const double *cam_to_world, int64_t capacity, double *xyz, float *normal,
                    uint8_t *grey, int64_t *count)
{
    if (!c) return VSM_EINVAL;
    if (!ids || n <= 0 || n > 64) return fail(c, VSM_EINVAL, "1..64 images required, got %d", n);
    if (depth_kind != 0 && depth_kind != 1) return fail(c, VSM_EINVAL, "depth_kind must be 0 or 1");
    if (!params || !cam_to_world || !count || capacity < 0 || (capacity > 0 && !xyz))
        return fail(c, VSM_EINVAL, "bad params / cam_to_world / outputs");
    if (params->min_num_pixels < 1 || !(params->max_reproj_error >= 0.f) ||
        !(params->max_depth_error >= 0.f) || !(params->max_normal_error_deg >= 0.f) ||
        !(params->max_normal_error_deg <= 180.f))
        return fail(c, VSM_EINVAL, "invalid fusion thresholds");
    std::vector<Fuse3dView> hv(n);
    long long max_hw = 0;
    for (int v = 0; v < n; ++v) {
        auto it = c->images.find(ids[v]);
        if (it == c->images.end()) return fail(c, VSM_ENOENT, "unknown image %d", ids[v]);
        const ImageRec &r = it->second;
        if (!r.depth[depth_kind] || !r.normal[depth_kind])
            return fail(c, VSM_EINVAL, "image %d has no %s depth + normal maps", ids[v],
                        depth_kind ? "geometric" : "photometric");
        Fuse3dView &f = hv[v];
        f.depth = r.depth[depth_kind];
        f.normal = r.normal[depth_kind];
        f.gray = r.gray;
        f.W = r.W;
        f.H = r.H;
        memcpy(f.P, r.P4, sizeof f.P);
        memcpy(f.Pinv, inv_proj_mats ? inv_proj_mats + 16 * v : r.Pinv, sizeof f.Pinv);
        for (int k = 0; k < 9; ++k) f.R[k] = (float)cam_to_world[9 * v + k];
        max_hw = std::max(max_hw, (long long)r.W * r.H);
    }
    Fuse3dConst k;
    k.n = n;
    k.min_num = params->min_num_pixels;
    k.max_depth_err = (double)params->max_depth_error;
    k.max_reproj = (double)params->max_reproj_error;
    k.cos_max_normal = (float)std::cos((double)params->max_normal_error_deg * M_PI / 180.0);
    HIPCHK(c, hipSetDevice(c->device));
    Fuse3dView *dviews = nullptr;
    FusedPoint *dense = nullptr, *sel = nullptr;
    int *flag = nullptr, *nsel = nullptr;
    double *dxyz = nullptr;
    float *dnrm = nullptr;
    uint8_t *dgray = nullptr, *temp = nullptr;
    size_t temp_bytes = 0;
    HIPCHK(c, compact_fused(nullptr, nullptr, max_hw, nullptr, temp_bytes, nullptr, nullptr, c->stream));
    int rc = dalloc(c, &dviews, n);
    if (!rc) rc = dalloc(c, &dense, max_hw);
    if (!rc) rc = dalloc(c, &sel, max_hw);
    if (!rc) rc = dalloc(c, &flag, max_hw);
    if (!rc) rc = dalloc(c, &nsel, 1);
    if (!rc) rc = dalloc(c, &dxyz, 3 * (size_t)max_hw);
    if (!rc) rc = dalloc(c, &dnrm, 3 * (size_t)max_hw);
    if (!rc) rc = dalloc(c, &dgray, max_hw);
    if (!rc) rc = dalloc(c, &temp, std::max<size_t>(temp_bytes, 1));
    auto cleanup = [&]() {
        dfree(c, dviews); dfree(c, dense); dfree(c, sel); dfree(c, flag); dfree(c, nsel);
        dfree(c, dxyz); dfree(c, dnrm); dfree(c, dgray); dfree(c, temp);
    };
    if (rc) { cleanup(); return rc; }
#define F3CHK(expr)                                                                             \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            cleanup();                                                                          \
            return fail(c, VSM_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));        \
        }                                                                                       \
    } while (0)
    F3CHK(hipMemcpyAsync(dviews, hv.data(), n * sizeof(Fuse3dView), hipMemcpyHostToDevice, c->stream));
    int64_t total = 0;
    for (int v = 0; v < n; ++v) {
        const long long hw = (long long)hv[v].W * hv[v].H;
        F3CHK(launch_fuse3d(dviews, v, hv[v].W, hv[v].H, k, dense, flag, c->stream));
        size_t tb = temp_bytes;
        F3CHK(compact_fused(dense, flag, hw, temp, tb, sel, nsel, c->stream));
        int ns = 0;
        F3CHK(hipMemcpyAsync(&ns, nsel, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        F3CHK(hipStreamSynchronize(c->stream));
        const int64_t room = std::max<int64_t>(0, capacity - total);
        const int64_t put = std::min<int64_t>(room, ns);
        if (put > 0) {
            F3CHK(launch_unpack_fused(sel, put, dxyz, dnrm, dgray, c->stream));
            F3CHK(hipMemcpyAsync(xyz + 3 * total, dxyz, put * 24, hipMemcpyDeviceToHost, c->stream));
            if (normal)
                F3CHK(hipMemcpyAsync(normal + 3 * total, dnrm, put * 12, hipMemcpyDeviceToHost, c->stream));
            if (grey) F3CHK(hipMemcpyAsync(grey + total, dgray, put, hipMemcpyDeviceToHost, c->stream));
            F3CHK(hipStreamSynchronize(c->stream));
        }
        total += ns;
    }
#undef F3CHK
    cleanup();
    *count = total;
    return VSM_OK;
}

int vsm_proj_to_grid(vsm_ctx *c, const double *points, int64_t n, const vsm_dsm_grid *grid,
                     double *dsm_out)
{
    if (!c) return VSM_EINVAL;
    int rc = check_grid(c, grid);
    if (rc) return rc;
    if ((n > 0 && !points) || !dsm_out || n < 0) return fail(c, VSM_EINVAL, "bad points/output");
    HIPCHK(c, hipSetDevice(c->device));
    const DsmGridDev g = grid_dev(*grid);
    const long long ncell = (long long)g.e_size * g.n_size;
    unsigned long long *cells = nullptr;
    double *pts = nullptr, *out = nullptr;
    rc = dalloc(c, &cells, ncell);
    if (!rc) rc = dalloc(c, &out, ncell);
    if (!rc && n > 0) rc = dalloc(c, &pts, 3 * (size_t)n);
    if (rc) { dfree(c, cells); dfree(c, out); return rc; }
    if (n > 0) HIPCHK(c, hipMemcpyAsync(pts, points, 24 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_fill_u64n(cells, ncell, 0ull, c->stream));
    HIPCHK(c, launch_rasterize_points(pts, n, g, cells, c->stream));
    HIPCHK(c, launch_decode_fill_f64(cells, g.e_size, g.n_size, out, c->stream));
    HIPCHK(c, hipMemcpyAsync(dsm_out, out, ncell * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c, cells); dfree(c, out); dfree(c, pts);
    return VSM_OK;
}

int vsm_proj_to_grid_noalloc(vsm_ctx *c, const double *points_dev, int64_t n, const vsm_dsm_grid *grid,
                             double *dsm_dev, void *stream)
{
    if (!c) return VSM_EINVAL;
    const int rc = check_grid(c, grid);
    if (rc) return rc;
    if (n < 0 || (n > 0 && !points_dev) || !dsm_dev) return fail(c, VSM_EINVAL, "bad points/output");
    const long long ncell = (long long)grid->e_size * grid->n_size;
    if (ncell > kNoAllocMaxCells)
        return fail(c, VSM_EINVAL, "grid of %lld cells exceeds the %lld cells of the allocation-free "
                    "scratch", ncell, (long long)kNoAllocMaxCells);
    HIPCHK(c, hipSetDevice(c->device));
    const hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    HIPCHK(c, proj_to_grid_noalloc(c->device, points_dev, n, grid_dev(*grid), dsm_dev, st));
    return VSM_OK;
}

// ---------------------------------------------------------------------------
// Geodetic frame (coordinate_system.py:41-51, lib/latlon_utm_converter.py:39-52)
// ---------------------------------------------------------------------------
int vsm_geo_frame_init(const vsm_geo_params *p, vsm_geo_frame *out)
{
    if (!p || !out) return VSM_EINVAL;
    if (p->zone_number < 1 || p->zone_number > 60 || !(std::fabs(p->lat0_deg) <= 84.0) ||
        !(std::fabs(p->lon0_deg) <= 180.0) || !std::isfinite(p->alt0))
        return VSM_EINVAL;
    const double deg = M_PI / 180.0;
    const double a = 6378137.0, f = 1.0 / 298.257223563;
    const double e2 = f * (2.0 - f), b = a * (1.0 - f), ep2 = e2 / (1.0 - e2);
    const double lat0 = p->lat0_deg * deg, lon0 = p->lon0_deg * deg;
    const double sf = std::sin(lat0), cf = std::cos(lat0), sl = std::sin(lon0), cl = std::cos(lon0);
    // geodetic -> ECEF of the origin
    const double N0 = a / std::sqrt(1.0 - e2 * sf * sf);
    vsm_geo_frame g{};
    g.origin[0] = (N0 + p->alt0) * cf * cl;
    g.origin[1] = (N0 + p->alt0) * cf * sl;
    g.origin[2] = (N0 * (1.0 - e2) + p->alt0) * sf;
    // columns: east, north, up unit vectors in ECEF
    const double R[9] = {-sl, -sf * cl, cf * cl,
                         cl,  -sf * sl, cf * sl,
                         0.0, cf,       sf};
    memcpy(g.rot, R, sizeof R);
    const double lam0 = (6.0 * p->zone_number - 183.0) * deg;
    g.cm_cos = std::cos(lam0);
    g.cm_sin = std::sin(lam0);
    g.a = a;
    g.e2 = e2;
    g.e = std::sqrt(e2);
    g.e2a = e2 * a;
    g.ep2b = ep2 * b;
    g.one_minus_f = 1.0 - f;
    // Krueger series to order n^6 (Karney 2011, eq. 35)
    const double n = f / (2.0 - f), n2 = n * n, n3 = n2 * n, n4 = n3 * n, n5 = n4 * n, n6 = n5 * n;
    const double A = a / (1.0 + n) * (1.0 + n2 / 4.0 + n4 / 64.0 + n6 / 256.0);
    g.k0A = 0.9996 * A;
    g.alpha[0] = n / 2.0 - 2.0 * n2 / 3.0 + 5.0 * n3 / 16.0 + 41.0 * n4 / 180.0 -
                 127.0 * n5 / 288.0 + 7891.0 * n6 / 37800.0;
    g.alpha[1] = 13.0 * n2 / 48.0 - 3.0 * n3 / 5.0 + 557.0 * n4 / 1440.0 + 281.0 * n5 / 630.0 -
                 1983433.0 * n6 / 1935360.0;
    g.alpha[2] = 61.0 * n3 / 240.0 - 103.0 * n4 / 140.0 + 15061.0 * n5 / 26880.0 +
                 167603.0 * n6 / 181440.0;
    g.alpha[3] = 49561.0 * n4 / 161280.0 - 179.0 * n5 / 168.0 + 6601661.0 * n6 / 7257600.0;
    g.alpha[4] = 34729.0 * n5 / 80640.0 - 3418889.0 * n6 / 1995840.0;
    g.alpha[5] = 212378941.0 * n6 / 319334400.0;
    g.false_easting = 500000.0;
    g.false_northing = p->south ? 10000000.0 : 0.0;
    g.max_q = 0.25;
    *out = g;
    return VSM_OK;
}

int vsm_set_geo_frame(vsm_ctx *c, const vsm_geo_frame *frame)
{
    if (!c) return VSM_EINVAL;
    if (!frame) {
        c->has_geo = false;
        return VSM_OK;
    }
    const double *v = reinterpret_cast<const double *>(frame);
    for (size_t i = 0; i < sizeof(vsm_geo_frame) / sizeof(double); ++i)
        if (!std::isfinite(v[i])) return fail(c, VSM_EINVAL, "non-finite geo frame constant %zu", i);
    if (!(frame->a > 0.0) || !(frame->max_q > 0.0) || !(frame->max_q <= 0.25))
        return fail(c, VSM_EINVAL, "geo frame: bad ellipsoid or max_q (0 < max_q <= 0.25)");
    c->geo = *frame;
    c->has_geo = true;
    return VSM_OK;
}

int vsm_enu_to_utm(vsm_ctx *c, const double *enu, int64_t n, double *out)
{
    if (!c) return VSM_EINVAL;
    if (!c->has_geo) return fail(c, VSM_EINVAL, "no geo frame set");
    if (n < 0 || (n > 0 && (!enu || !out))) return fail(c, VSM_EINVAL, "bad points/output");
    if (n == 0) return VSM_OK;
    HIPCHK(c, hipSetDevice(c->device));
    double *buf = nullptr;
    int rc = dalloc(c, &buf, 6 * (size_t)n);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(buf, enu, 24 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_enu_to_utm(buf, n, c->geo, buf + 3 * (size_t)n, c->stream));
    HIPCHK(c, hipMemcpyAsync(out, buf + 3 * (size_t)n, 24 * (size_t)n, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c, buf);
    return VSM_OK;
}

int vsm_median_blur3(vsm_ctx *c, const float *in, int32_t e_size, int32_t n_size, float *out)
{
    if (!c) return VSM_EINVAL;
    if (!in || !out || e_size <= 0 || n_size <= 0 || (int64_t)e_size * n_size > (1ll << 31))
        return fail(c, VSM_EINVAL, "bad raster %dx%d", e_size, n_size);
    HIPCHK(c, hipSetDevice(c->device));
    const size_t nc = (size_t)e_size * n_size;
    float *buf = nullptr;
    int rc = dalloc(c, &buf, 2 * nc);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(buf, in, nc * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_median3(buf, e_size, n_size, buf + nc, c->stream));
    HIPCHK(c, hipMemcpyAsync(out, buf + nc, nc * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c, buf);
    return VSM_OK;
}

// ---------------------------------------------------------------------------
// Serving: continuous batching of per-view solves, photometric and (once a
// view's photometric result and its sources' photometric depth maps are
// resident) geometric consistency.
//
// Results are those of running the requests one at a time in ticket order: a
// geometric request waits for every earlier photometric request of its
// reference or of one of its sources, a photometric request for every earlier
// request that reads or writes its image's photometric maps.  Within those
// constraints requests share launches, and a request's result does not depend
// on its co-runners (tiles are independent within a half-step and the random
// streams are keyed by view and pixel).
// ---------------------------------------------------------------------------

// Scratch the loop allocates for a request besides its state: its share of
// the served tile list (reallocated at twice the size needed), the filter's
// and the init's tile lists and descriptors, and block rounding.
static size_t serve_scratch_bytes(int W, int H)
{
    // tile counts of the narrowest half-step tiles (16 x 32) bound every width's
    const size_t ntiles = (size_t)((W + 15) / 16) * ((H + 31) / 32) + (size_t)((W + 63) / 64) * ((H + 7) / 8);
    return 4 * ntiles * sizeof(Tile) + 2 * sizeof(ViewDesc) + 16 * kAllocAlign;
}

// Device bytes a request needs while active: state, outputs (depth, normal,
// cost, consistency mask), the image maps its result is kept in when the image
// has none yet, the packed depth textures of its sources (geometric pass) that
// are not resident, and its scratch.  worst: as if nothing were resident.
static size_t serve_footprint(vsm_ctx *c, const ServeReq &r, bool worst)
{
    const ImageRec &ref = c->images.at(r.ref);
    const size_t hw = (size_t)ref.W * ref.H;
    size_t b = state_bytes(ref.W, ref.H, (int)r.srcs.size()) + hw * 24 + serve_scratch_bytes(ref.W, ref.H);
    if (worst || !ref.depth[r.pass]) b += hw * 4;
    if (worst || !ref.normal[r.pass]) b += hw * 12;
    if (r.pass == 1)
        for (int s : r.srcs) {
            const ImageRec &S = c->images.at(s);
            if (worst || !S.depth4) b += (size_t)S.W * S.H * sizeof(float4);
        }
    return b;
}

static bool has_id(const std::vector<int> &v, int id)
{
    return std::find(v.begin(), v.end(), id) != v.end();
}

// Whether the pending request e (earlier or later) must finish before r starts.
static bool serve_blocks(const ServeReq &e, const ServeReq &r)
{
    if (e.ticket >= r.ticket) return false;
    if (r.pass == 0)          // r rewrites its image's photometric maps
        return e.ref == r.ref || (e.pass == 1 && has_id(e.srcs, r.ref));
    if (e.pass == 0)          // r reads e's result
        return e.ref == r.ref || has_id(r.srcs, e.ref);
    return e.ref == r.ref;    // both write the geometric maps
}

// 1: r can start now; 0: it waits for an earlier request; negative: it can
// never start (code and message set).
static int serve_ready(vsm_ctx *c, const ServeReq &r)
{
    for (const ServeReq *e : c->active)
        if (serve_blocks(*e, r)) return 0;
    for (const auto &e : c->queue)
        if (serve_blocks(*e, r)) return 0;
    auto ri = c->images.find(r.ref);
    if (ri == c->images.end()) return fail(c, VSM_ENOENT, "unknown reference image %d", r.ref);
    for (int s : r.srcs)
        if (c->images.find(s) == c->images.end()) return fail(c, VSM_ENOENT, "unknown source image %d", s);
    if (r.pass == 1) {
        if (!ri->second.depth[0] || !ri->second.normal[0])
            return fail(c, VSM_EINVAL, "geometric request for image %d: it has no photometric depth and "
                        "normal maps", r.ref);
        for (int s : r.srcs)
            if (!c->images.at(s).depth[0])
                return fail(c, VSM_EINVAL, "geometric request for image %d: source %d has no photometric "
                            "depth map", r.ref, s);
    }
    return 1;
}

static void serve_drop_keep(vsm_ctx *c, ServeReq &r)
{
    dfree(c, r.keep_depth);
    dfree(c, r.keep_normal);
    r.keep_depth = nullptr;
    r.keep_normal = nullptr;
}

// State, outputs and initial hypotheses of an admitted request; on failure
// everything it allocated is released again.
// The request's half-step tiles at the context's tile width on the device,
// its slot as their view index.
// The host copy stays with the request until its tiles are rebuilt or it
// retires, so the asynchronous upload needs no wait here.
static int serve_build_tiles(vsm_ctx *c, ServeReq &r)
{
    const int tw = c->params.tile_width;
    std::vector<Tile> tl;
    tiles_for(r.vs.desc, r.slot, tl, tw);
    Tile *dt = nullptr;
    int rc = dalloc(c, &dt, tl.size());
    if (rc) return rc;
    const hipError_t e = hipMemcpyAsync(dt, tl.data(), tl.size() * sizeof(Tile), hipMemcpyHostToDevice,
                                        c->stream);
    if (e != hipSuccess) {
        dfree(c, dt);
        return fail(c, VSM_EDEVICE, "request tiles: %s", hipGetErrorString(e));
    }
    if (r.dtiles) {
        // the previous tiles may still be read by queued work
        HIPCHK(c, hipStreamSynchronize(c->stream));
        dfree(c, r.dtiles);
    }
    r.dtiles = dt;
    r.htiles.swap(tl);
    r.ntiles = (int)r.htiles.size();
    r.tile_w = tw;
    c->serve_dirty = true;
    return VSM_OK;
}

static void serve_release_slot(vsm_ctx *c, ServeReq &r)
{
    if (r.slot >= 0) c->serve_slots[r.slot] = nullptr;
    r.slot = -1;
    dfree(c, r.dtiles);
    r.dtiles = nullptr;
    r.ntiles = 0;
    std::vector<Tile>().swap(r.htiles);
    c->serve_dirty = true;
}

// State, outputs, descriptor slot, tiles and initial hypotheses of an
// admitted request; on failure everything it took is released again.
static int serve_start(vsm_ctx *c, ServeReq &r)
{
    vsm_view_req q;
    memset(&q, 0, sizeof q);
    q.ref_id = r.ref;
    q.n_src = (int)r.srcs.size();
    q.src_ids = r.srcs.data();
    q.depth_min = r.dmin;
    q.depth_max = r.dmax;
    int rc = prepare_view(c, q, r.pass, (uint32_t)r.ref, r.vs);
    if (rc) return rc;
    const ImageRec &img = c->images.at(r.ref);
    const size_t hw = (size_t)r.vs.desc.W * r.vs.desc.H;
    for (int k = 0; k < (int)c->serve_slots.size() && r.slot < 0; ++k)
        if (!c->serve_slots[k]) r.slot = k;
    if (r.slot < 0) rc = fail(c, VSM_EINVAL, "no free serving slot");
    if (!rc) {
        c->serve_slots[r.slot] = &r;
        rc = serve_build_tiles(c, r);
    }
    if (!rc) rc = dalloc(c, &r.out_depth, hw * 6);
    if (!rc && !img.depth[r.pass]) rc = dalloc(c, &r.keep_depth, hw);
    if (!rc && !img.normal[r.pass]) rc = dalloc(c, &r.keep_normal, hw * 3);
    if (!rc) {
        const PMConst pc = make_const(c->params);
        ViewDesc *dv = c->serve_views + r.slot;
        r.vs.desc.step = 0;
        hipError_t e = hipMemcpyAsync(dv, &r.vs.desc, sizeof(ViewDesc), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess && r.pass == 1)     // start from the photometric depth and normals
            e = launch_load_hyp(dv, 0, (int)hw, img.depth[0], img.normal[0], pc.cos_max_normal, c->stream);
        if (e == hipSuccess)
            e = launch_init(c->serve_views, r.dtiles, r.ntiles, pc, r.pass == 0 ? 1 : 0, r.tile_w, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(c, VSM_EDEVICE, "request start failed: %s", hipGetErrorString(e));
    }
    if (rc) {
        dfree(c, r.out_depth);
        r.out_depth = nullptr;
        serve_drop_keep(c, r);
        serve_release_slot(c, r);
        release_view(c, r.vs);
        return rc;
    }
    r.out_normal = r.out_depth + hw;
    r.out_cost = r.out_depth + 4 * hw;
    r.out_mask = reinterpret_cast<uint32_t *>(r.out_depth + 5 * hw);
    r.host_step = 0;
    return VSM_OK;
}

static void serve_fail_req(vsm_ctx *c, std::unique_ptr<ServeReq> owned, int code)
{
    owned->done = true;
    owned->error = code;
    owned->errmsg = c->err;
    c->reqs[owned->ticket] = std::move(owned);
}

// Admit, in ticket order, every request that can start and fits the budget.
// A request that waits for an earlier one is passed over; one that is ready
// but does not fit stops admission (first come, first served for memory),
// unless nothing is active: then nothing will free memory and it fails.
static int serve_admit(vsm_ctx *c)
{
    auto it = c->queue.begin();
    while (it != c->queue.end() && (int)c->active.size() < kServeMaxViews) {
        ServeReq &r = **it;
        const int ready = serve_ready(c, r);
        if (ready < 0) {
            serve_fail_req(c, std::move(*it), ready);
            it = c->queue.erase(it);
            continue;
        }
        if (ready == 0) {
            ++it;
            continue;
        }
        const size_t fp = serve_footprint(c, r, false);
        if (c->budget && c->in_use + fp > c->budget) {
            if (!c->active.empty()) break;
            fail(c, VSM_ENOMEM, "request for image %d needs %zu bytes; %llu of the %llu-byte budget "
                 "are held by resident images and maps", r.ref, fp, (unsigned long long)c->in_use,
                 (unsigned long long)c->budget);
            serve_fail_req(c, std::move(*it), VSM_ENOMEM);
            it = c->queue.erase(it);
            continue;
        }
        const int rc = serve_start(c, r);
        if (rc == VSM_ENOMEM && !c->active.empty()) break;     // memory frees up later
        if (rc) {
            if (rc == VSM_EDEVICE) return rc;
            serve_fail_req(c, std::move(*it), rc);
            it = c->queue.erase(it);
            continue;
        }
        c->active.push_back(&r);
        c->reqs[r.ticket] = std::move(*it);
        it = c->queue.erase(it);
    }
    return VSM_OK;
}

// Filter (when the request asks for it) and outputs of a request whose last
// half-step has run; dv is its descriptor slot in the served view array.
static int serve_finish(vsm_ctx *c, ServeReq &r, ViewDesc *dv)
{
    ImageRec &img = c->images.at(r.ref);
    const size_t hw = (size_t)r.vs.desc.W * r.vs.desc.H;
    const PMConst pc = make_const(c->params);
    r.vs.desc.step = r.vs.desc.num_steps;
    HIPCHK(c, hipMemcpyAsync(dv, &r.vs.desc, sizeof(ViewDesc), hipMemcpyHostToDevice, c->stream));
    if (r.filter && r.tile_w == kTileW && r.dtiles) {
        // the request's own half-step tiles have the filter's shape
        HIPCHK(c, launch_filter(r.vs.desc.n_src, c->serve_views, r.dtiles, r.ntiles, pc, r.out_mask,
                                c->stream));
    } else if (r.filter) {
        std::vector<Tile> tl;
        tiles_for(r.vs.desc, 0, tl);
        Tile *dt = nullptr;
        const int rc = dalloc(c, &dt, tl.size());
        if (rc) return rc;
        hipError_t e = hipMemcpyAsync(dt, tl.data(), tl.size() * sizeof(Tile), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = launch_filter(r.vs.desc.n_src, dv, dt, (int)tl.size(), pc, r.out_mask, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        dfree(c, dt);
        if (e != hipSuccess) return fail(c, VSM_EDEVICE, "filter failed: %s", hipGetErrorString(e));
    } else {
        HIPCHK(c, launch_fill_u32(r.out_mask, (long long)hw, 0u, c->stream));
    }
    // the result also becomes the image's depth and normal map of this pass
    // on the device (fusion, later geometric passes), as after vsm_solve
    if (!img.depth[r.pass]) { img.depth[r.pass] = r.keep_depth; r.keep_depth = nullptr; }
    if (!img.normal[r.pass]) { img.normal[r.pass] = r.keep_normal; r.keep_normal = nullptr; }
    HIPCHK(c, launch_finalize(dv, 0, (int)hw, r.out_depth, r.out_normal, r.out_cost, img.depth[r.pass],
                              c->stream));
    HIPCHK(c, hipMemcpyAsync(img.normal[r.pass], r.out_normal, hw * 12, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (r.pass == 0) ++img.depth_ver;
    return VSM_OK;
}

static int serve_submit(vsm_ctx *c, const vsm_view_req *req, int pass, int filter, uint64_t *ticket)
{
    if (!c || !req || !ticket) return VSM_EINVAL;
    if (pass != 0 && pass != 1) return fail(c, VSM_EINVAL, "pass must be 0 or 1");
    if (filter < -1 || filter > 1) return fail(c, VSM_EINVAL, "filter must be -1, 0 or 1");
    if (req->n_src < 1 || req->n_src > VSM_MAX_SRC || !req->src_ids)
        return fail(c, VSM_EINVAL, "n_src %d outside [1, %d]", req->n_src, VSM_MAX_SRC);
    if (c->images.find(req->ref_id) == c->images.end())
        return fail(c, VSM_ENOENT, "unknown reference image %d", req->ref_id);
    for (int s = 0; s < req->n_src; ++s)
        if (c->images.find(req->src_ids[s]) == c->images.end())
            return fail(c, VSM_ENOENT, "unknown source image %d", req->src_ids[s]);
    std::lock_guard<std::mutex> lk(c->serve_mu);
    HIPCHK(c, hipSetDevice(c->device));
    // resident serving scratch is allocated before any admission decision
    int rc = VSM_OK;
    if (!c->serve_views && (rc = dalloc(c, &c->serve_views, kServeMaxViews))) return rc;
    c->serve_slots.resize(kServeMaxViews, nullptr);
    if (!tile_width_supported(c->params.tile_width, req->n_src))
        return fail(c, VSM_EINVAL, "tile_width %d is not available for %d sources", c->params.tile_width,
                    req->n_src);
    if ((rc = ensure_c0(c, req->n_src, c->params.tile_width))) return rc;
    auto r = std::make_unique<ServeReq>();
    r->ref = req->ref_id;
    r->pass = pass;
    r->filter = filter < 0 ? (c->params.filter ? 1 : 0) : filter;
    r->srcs.assign(req->src_ids, req->src_ids + req->n_src);
    r->dmin = req->depth_min;
    r->dmax = req->depth_max;
    const size_t worst = serve_footprint(c, *r, true);
    if (c->budget && worst > c->budget)
        return fail(c, VSM_ENOMEM, "request footprint %zu exceeds the HBM budget", worst);
    r->ticket = c->next_ticket++;
    *ticket = r->ticket;
    c->queue.push_back(std::move(r));
    return VSM_OK;
}

int vsm_submit(vsm_ctx *c, const vsm_view_req *req, uint64_t *ticket)
{
    return serve_submit(c, req, 0, -1, ticket);
}

int vsm_submit_pass(vsm_ctx *c, const vsm_view_req *req, int32_t pass, int32_t filter, uint64_t *ticket)
{
    return serve_submit(c, req, pass, filter, ticket);
}

int vsm_serve_step(vsm_ctx *c, int32_t *n_active, int32_t *n_queued)
{
    if (!c) return VSM_EINVAL;
    std::lock_guard<std::mutex> lk(c->serve_mu);
    HIPCHK(c, hipSetDevice(c->device));
    int rc = serve_admit(c);
    if (rc) return rc;
    const int na = (int)c->active.size();
    if (na > 0) {
        // The descriptors and tiles stay on the device while a request is
        // active: a step re-concatenates the tile lists only when the active
        // list (or the tile width) changed, and the device step counters
        // advance after each half-step (k_advance), in step with host_step.
        int maxs = 1, nslots = 0;
        for (ServeReq *r : c->active) {
            maxs = std::max(maxs, r->vs.desc.n_src);
            nslots = std::max(nslots, r->slot + 1);
            if (r->tile_w != c->params.tile_width && (rc = serve_build_tiles(c, *r))) return rc;
        }
        if (!tile_width_supported(c->params.tile_width, maxs))
            return fail(c, VSM_EINVAL, "tile_width %d is not available for %d sources", c->params.tile_width,
                        maxs);
        if (c->serve_dirty) {
            long long total = 0;
            for (ServeReq *r : c->active) total += r->ntiles;
            if (total > c->serve_cap_tiles) {
                dfree(c, c->serve_tiles);
                c->serve_tiles = nullptr;
                c->serve_cap_tiles = 0;
                if ((rc = dalloc(c, &c->serve_tiles, (size_t)total * 2))) return rc;
                c->serve_cap_tiles = (int)(total * 2);
            }
            long long off = 0;
            for (ServeReq *r : c->active) {
                HIPCHK(c, hipMemcpyAsync(c->serve_tiles + off, r->dtiles, r->ntiles * sizeof(Tile),
                                         hipMemcpyDeviceToDevice, c->stream));
                off += r->ntiles;
            }
            c->serve_ntiles = (int)total;
            c->serve_dirty = false;
        }
        const PMConst pc = make_const(c->params);
        if ((rc = record_halfstep(c, maxs, c->params.tile_width, c->serve_views, c->serve_tiles,
                                  c->serve_ntiles, pc, c->timing)))
            return rc;
        HIPCHK(c, launch_advance(c->serve_views, nslots, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        // Every active request has advanced one half-step.  The finished ones
        // leave the active list before any fallible work on them, so a
        // failure below affects only the request it belongs to.
        std::vector<std::pair<ServeReq *, int>> retire;
        std::vector<ServeReq *> still;
        for (int v = 0; v < na; ++v) {
            ServeReq &r = *c->active[v];
            r.host_step += 1;
            if (r.host_step < r.vs.desc.num_steps) still.push_back(&r);
            else retire.emplace_back(&r, r.slot);
        }
        if (!retire.empty()) c->serve_dirty = true;
        c->active.swap(still);
        int dev_err = VSM_OK;
        for (auto &rv : retire) {
            ServeReq &r = *rv.first;
            const int e = serve_finish(c, r, c->serve_views + rv.second);
            serve_release_slot(c, r);
            release_view(c, r.vs);
            serve_drop_keep(c, r);
            r.done = true;
            if (e) {
                r.error = e;
                r.errmsg = c->err;
                dfree(c, r.out_depth);
                r.out_depth = nullptr;
                if (e == VSM_EDEVICE && !dev_err) dev_err = e;
            }
        }
        rc = read_counters(c, pc);
        if (rc) return rc;
        if (dev_err) return dev_err;
    }
    if (n_active) *n_active = (int)c->active.size();
    if (n_queued) *n_queued = (int)c->queue.size();
    return VSM_OK;
}

int vsm_poll(vsm_ctx *c, uint64_t ticket, vsm_view_out *out, int32_t *done)
{
    if (!c || !done) return VSM_EINVAL;
    std::lock_guard<std::mutex> lk(c->serve_mu);
    auto it = c->reqs.find(ticket);
    if (it == c->reqs.end()) {
        for (auto &q : c->queue)
            if (q->ticket == ticket) { *done = 0; return VSM_OK; }
        return fail(c, VSM_ENOENT, "unknown ticket %llu", (unsigned long long)ticket);
    }
    ServeReq &r = *it->second;
    if (!r.done) { *done = 0; return VSM_OK; }
    *done = 1;
    if (r.error) {
        c->err = r.errmsg;
        int e = r.error;
        c->reqs.erase(it);
        return e;
    }
    const size_t hw = (size_t)r.vs.desc.W * r.vs.desc.H;
    if (out && out->depth) HIPCHK(c, hipMemcpyAsync(out->depth, r.out_depth, hw * 4, hipMemcpyDeviceToHost, c->stream));
    if (out && out->normal) HIPCHK(c, hipMemcpyAsync(out->normal, r.out_normal, hw * 12, hipMemcpyDeviceToHost, c->stream));
    if (out && out->cost) HIPCHK(c, hipMemcpyAsync(out->cost, r.out_cost, hw * 4, hipMemcpyDeviceToHost, c->stream));
    if (out && out->consistency)
        HIPCHK(c, hipMemcpyAsync(out->consistency, r.out_mask, hw * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c, r.out_depth);
    c->reqs.erase(it);
    return VSM_OK;
}

int vsm_set_hbm_budget(vsm_ctx *c, uint64_t hbm_budget_bytes)
{
    if (!c) return VSM_EINVAL;
    std::lock_guard<std::mutex> lk(c->serve_mu);
    c->budget = hbm_budget_bytes;
    if (c->budget) release_cached(c, 0);
    return VSM_OK;
}

int vsm_get_stats(vsm_ctx *c, vsm_stats *o)
{
    if (!c || !o) return VSM_EINVAL;
    memset(o, 0, sizeof *o);
    o->halfstep_ms = c->halfstep_ms;
    o->halfstep_launches = c->halfstep_launches;
    o->ncc_evals = c->acc_ncc;
    o->geo_evals = c->acc_geo;
    o->taps = c->acc_taps;
    o->fuse_ms = c->fuse_ms;
    o->fuse_launches = c->fuse_launches;
    o->fuse_pixels = c->fuse_pixels;
    o->hbm_bytes_in_use = c->in_use;
    o->hbm_bytes_peak = c->peak;
    return VSM_OK;
}

int vsm_reset_stats(vsm_ctx *c)
{
    if (!c) return VSM_EINVAL;
    c->halfstep_ms = 0.0;
    c->halfstep_launches = 0;
    c->acc_ncc = c->acc_geo = c->acc_taps = 0;
    c->fuse_ms = 0.0;
    c->fuse_launches = c->fuse_pixels = 0;
    c->peak = c->in_use;
    return VSM_OK;
}

int vsm_set_timing(vsm_ctx *c, int32_t enabled)
{
    if (!c) return VSM_EINVAL;
    c->timing = enabled != 0;
    return VSM_OK;
}

int vsm_synchronize(vsm_ctx *c)
{
    if (!c) return VSM_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VSM_OK;
}

}  // extern "C"
