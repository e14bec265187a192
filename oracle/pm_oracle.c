/*
 * pm_oracle.c -- FP64 CPU restatement of the PatchMatch MVS semantics.
 *
 * TEST INFRASTRUCTURE (the oracle).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, and only as the checker or
 * the timed CPU baseline; the product path never links or calls it.
 *
 * Provenance.  The reference runs PatchMatch in an external, un-vendored
 * "adapted COLMAP" binary (reference README.md:6, colmap_mvs_commands.py:40-82),
 * so the per-pixel arithmetic is not in the reference tree.  What the
 * reference fixes, and this file follows:
 *   - the projective point p = (col, row, 1, d) with the 4x4
 *     P4 = [K[R|t]; depth_min*(0,0,1,-z0)] (reparam_depth.py:104,136-141) and
 *     its inverse for back-projection (aggregate_2p5d_util.py:86-90);
 *   - the flags: window_radius, min_triangulation_angle 5.0, num_samples 15,
 *     geom_consistency_regularizer 0.3, geom_consistency_max_cost 3,
 *     filter_min_triangulation_angle 4.99, filter_min_ncc 0.1,
 *     filter_geom_consistency_max_cost 1, filter_min_num_consistent 2
 *     (colmap_mvs_commands.py:40-75);
 *   - the depth range file (reparam_depth.py:187-189) and d <= 0 as "no depth"
 *     (aggregate_2p5d_util.py:76).
 * The remaining semantics restate the public PatchMatch-MVS algorithm
 * (bilaterally weighted NCC, sampled view selection with triangulation- and
 * incident-angle priors, geometric forward-backward consistency, filter) with
 * this project's red-black checkerboard schedule.  PARITY UNPINNED against the
 * reference for these parts: no reference code, test or vector exists for
 * them.  They are pinned by closed forms and brute force in tests/.
 *
 * Semantics (shared with the HIP kernels):
 *   hypothesis h = (d, n): reparametrised depth at the pixel centre and a unit
 *   normal n (n2 > 0) of a plane in (col, row, d) space;
 *   plane depth at offset (dx,dy): d + a dx + b dy, a = -n0/n2, b = -n1/n2;
 *   source pixel of a tap: q = M_s (x+dx, y+dy, 1, d+a dx+b dy), dehomogenised;
 *   cost C_photo = 1 - NCC_w in [0,2], bilateral w = exp(-(dx^2+dy^2)/(2 ss^2)
 *   - (r_t - r_c)^2/(2 sc^2)) on [0,1] intensities; MAX_COST = 2 when the centre
 *   projects outside the source or either weighted variance < 1e-5;
 *   geometric pass: C = C_photo + lambda * min(e_fb, geom_max) with e_fb the
 *   forward-backward reprojection error through the source depth map.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAX_SRC 64
#define ORC_MAX_COST 2.0
#define ORC_MIN_VAR 1e-5
#define ORC_NHYP 8

typedef struct {
    int32_t window_radius;
    int32_t window_step;
    double sigma_spatial;          /* pixels */
    double sigma_color;            /* on [0,1] intensities */
    double ncc_sigma;
    int32_t num_samples;
    double min_tri_angle;          /* radians */
    double incident_angle_sigma;   /* radians */
    double geom_regularizer;
    double geom_max_cost;          /* pixels */
    double filter_min_ncc;
    double filter_min_tri_angle;   /* radians */
    double filter_geom_max_cost;   /* pixels */
    int32_t filter_min_num_consistent;
    double depth_min, depth_max;
    double perturbation;           /* fraction of the depth range */
    double normal_perturbation;    /* radians */
    double max_normal_angle;       /* radians, cone of random normals */
    double perturb_decay;          /* per iteration */
    int32_t n_far;                 /* candidates per direction */
    uint64_t seed;
    int32_t draw_group;            /* 0: source draws keyed by pixel; else the width (16, 32, 64)
                                      of the pixel groups, 128 / width rows tall, whose pixels
                                      share the draws' uniforms */
} orc_params;

typedef struct {
    int32_t width, height;
    const uint8_t *ref;
    int32_t n_src;
    const uint8_t *const *src;
    const int32_t *src_w, *src_h;
    const double *M;                /* n_src x 12, ref (x,y,1,d) -> src */
    const double *Mb;               /* n_src x 12, src (x,y,1,d) -> ref */
    const double *tri;              /* n_src, radians */
    const double *g;                /* n_src x 3, J v_s */
    const double *J;                /* 3x3 row-major */
    const float *const *src_depth;  /* n_src depth maps or NULL */
    uint32_t view_id;
} orc_view;

/* ---------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).                      */
/* ---------------------------------------------------------------------- */
void orc_philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline double u01(uint32_t x) { return (double)(x >> 8) * (1.0 / 16777216.0); }

static void rng_block(const orc_view *v, const orc_params *p, uint32_t pixel, uint32_t step,
                      uint32_t block, double u[4])
{
    uint32_t ctr[4] = {pixel, step, block, 0u};
    uint32_t key[2] = {(uint32_t)p->seed, (uint32_t)(p->seed >> 32) ^ v->view_id};
    uint32_t o[4];
    orc_philox(ctr, key, o);
    for (int i = 0; i < 4; ++i) u[i] = u01(o[i]);
}

static inline uint32_t step_code(int pass, int iter, int color)
{
    return ((uint32_t)pass << 24) | ((uint32_t)iter << 1) | (uint32_t)color;
}
static inline uint32_t init_code(int pass) { return ((uint32_t)pass << 24) | 0xFFFFFFu; }

/* ---------------------------------------------------------------------- */
static inline double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

static double bilinear_u8(const uint8_t *img, int w, int h, double x, double y)
{
    double x0f = floor(x), y0f = floor(y);
    int x0 = (int)x0f, y0 = (int)y0f;
    int x1 = x0 + 1 < w ? x0 + 1 : w - 1;
    int y1 = y0 + 1 < h ? y0 + 1 : h - 1;
    double fx = x - x0f, fy = y - y0f;
    double a = img[(size_t)y0 * w + x0], b = img[(size_t)y0 * w + x1];
    double c = img[(size_t)y1 * w + x0], d = img[(size_t)y1 * w + x1];
    double top = a + fx * (b - a);
    double bot = c + fx * (d - c);
    return top + fy * (bot - top);
}

static inline void proj(const double *M, double x, double y, double d, double q[3])
{
    q[0] = M[0] * x + M[1] * y + M[2] + M[3] * d;
    q[1] = M[4] * x + M[5] * y + M[6] + M[7] * d;
    q[2] = M[8] * x + M[9] * y + M[10] + M[11] * d;
}

/* Homogeneous source position of the window tap (dx, dy) of pixel (x, y)
 * under the plane of depth d and slopes (a, b): the plane depth at the tap is
 * d + a dx + b dy, projected through M. */
static inline void tap_proj(const double *M, int x, int y, double d, double a, double b, int dx,
                            int dy, double t[3])
{
    const double dd = d + a * dx + b * dy;
    proj(M, x + dx, y + dy, dd, t);
}

/* Source pixel (not clamped) of the window tap (dx, dy) of pixel (x, y) under
 * the hypothesis (d, n) against source s; returns 0 when the tap's homogeneous
 * depth is not positive.  Exported so the tap geometry can be checked against
 * the plane-induced homography of debuggers/sfm_debug_warping_subarea.py:45-57. */
int orc_tap_coord(const orc_view *v, int s, int x, int y, double d, const double n[3], int dx,
                  int dy, double out[2])
{
    const double a = -n[0] / n[2], b = -n[1] / n[2];
    double t[3];
    tap_proj(v->M + 12 * s, x, y, d, a, b, dx, dy, t);
    out[0] = t[0] / t[2];
    out[1] = t[1] / t[2];
    return t[2] > 0.0;
}

/* 1 - bilaterally weighted NCC between the reference window at (x,y) and its
 * image in source s under the plane (d, n). */
double orc_photo_cost(const orc_view *v, const orc_params *p, int x, int y, double d,
                      const double n[3], int s)
{
    const double *M = v->M + 12 * s;
    const int W = v->width, H = v->height;
    const int Ws = v->src_w[s], Hs = v->src_h[s];
    const uint8_t *src = v->src[s];
    const double a = -n[0] / n[2], b = -n[1] / n[2];
    double q[3];
    proj(M, x, y, d, q);
    if (!(q[2] > 0.0)) return ORC_MAX_COST;
    const double cx = q[0] / q[2], cy = q[1] / q[2];
    if (!(cx >= 0.0 && cx <= Ws - 1 && cy >= 0.0 && cy <= Hs - 1)) return ORC_MAX_COST;
    const double sc = bilinear_u8(src, Ws, Hs, cx, cy) / 255.0;
    const double rc = v->ref[(size_t)y * W + x] / 255.0;
    const int r = p->window_radius, st = p->window_step;
    const int nt = (2 * r) / st + 1;
    const double inv2ss = 1.0 / (2.0 * p->sigma_spatial * p->sigma_spatial);
    const double inv2sc = 1.0 / (2.0 * p->sigma_color * p->sigma_color);
    double sw = 0, sr = 0, srr = 0, ss = 0, sss = 0, srs = 0;
    for (int j = 0; j < nt; ++j) {
        const int dy = -r + j * st;
        const int ry = y + dy < 0 ? 0 : (y + dy >= H ? H - 1 : y + dy);
        for (int i = 0; i < nt; ++i) {
            const int dx = -r + i * st;
            const int rx = x + dx < 0 ? 0 : (x + dx >= W ? W - 1 : x + dx);
            const double rv = v->ref[(size_t)ry * W + rx] / 255.0 - rc;
            double t[3];
            tap_proj(M, x, y, d, a, b, dx, dy, t);
            const double px = clampd(t[0] / t[2], 0.0, Ws - 1);
            const double py = clampd(t[1] / t[2], 0.0, Hs - 1);
            const double sv = bilinear_u8(src, Ws, Hs, px, py) / 255.0 - sc;
            const double w = exp(-(double)(dx * dx + dy * dy) * inv2ss - rv * rv * inv2sc);
            sw += w;
            sr += w * rv;
            srr += w * rv * rv;
            ss += w * sv;
            sss += w * sv * sv;
            srs += w * rv * sv;
        }
    }
    const double mr = sr / sw, ms = ss / sw;
    const double var_r = srr / sw - mr * mr;
    const double var_s = sss / sw - ms * ms;
    if (var_r < ORC_MIN_VAR || var_s < ORC_MIN_VAR) return ORC_MAX_COST;
    const double cov = srs / sw - mr * ms;
    const double ncc = cov / sqrt(var_r * var_s);
    return clampd(1.0 - ncc, 0.0, ORC_MAX_COST);
}

/* Forward-backward reprojection error (pixels) through source s's depth map;
 * geom_max when the round trip is undefined. */
double orc_geo_error(const orc_view *v, const orc_params *p, int x, int y, double d, int s)
{
    const double emax = p->geom_max_cost;
    if (!v->src_depth || !v->src_depth[s]) return emax;
    const int Ws = v->src_w[s], Hs = v->src_h[s];
    double q[3];
    proj(v->M + 12 * s, x, y, d, q);
    if (!(q[2] > 0.0)) return emax;
    const double qx = q[0] / q[2], qy = q[1] / q[2];
    if (!(qx >= 0.0 && qx <= Ws - 1 && qy >= 0.0 && qy <= Hs - 1)) return emax;
    const float *D = v->src_depth[s];
    const double x0f = floor(qx), y0f = floor(qy);
    const int x0 = (int)x0f, y0 = (int)y0f;
    const int x1 = x0 + 1 < Ws ? x0 + 1 : Ws - 1, y1 = y0 + 1 < Hs ? y0 + 1 : Hs - 1;
    const double da = D[(size_t)y0 * Ws + x0], db = D[(size_t)y0 * Ws + x1];
    const double dc = D[(size_t)y1 * Ws + x0], de = D[(size_t)y1 * Ws + x1];
    if (!(da > 0 && db > 0 && dc > 0 && de > 0)) return emax;
    const double fx = qx - x0f, fy = qy - y0f;
    const double top = da + fx * (db - da), bot = dc + fx * (de - dc);
    const double ds = top + fy * (bot - top);
    double b[3];
    proj(v->Mb + 12 * s, qx, qy, ds, b);
    if (!(b[2] > 0.0)) return emax;
    const double ex = b[0] / b[2] - x, ey = b[1] / b[2] - y;
    return sqrt(ex * ex + ey * ey);
}

static double total_cost(const orc_view *v, const orc_params *p, int pass, int x, int y,
                         double d, const double n[3], int s, double photo)
{
    double c = photo;
    if (pass == 1) {
        double e = orc_geo_error(v, p, x, y, d, s);
        c += p->geom_regularizer * (e < p->geom_max_cost ? e : p->geom_max_cost);
    }
    return c;
}

/* Cost of fixed per-pixel hypotheses for every source: out[s*H*W + y*W + x]. */
void orc_cost_map(const orc_view *v, const orc_params *p, int pass, const double *depth,
                  const double *normal, double *out)
{
    const int W = v->width, H = v->height;
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            const double *n = normal + 3 * i;
            for (int s = 0; s < v->n_src; ++s) {
                double c = orc_photo_cost(v, p, x, y, depth[i], n, s);
                out[(size_t)s * W * H + i] = total_cost(v, p, pass, x, y, depth[i], n, s, c);
            }
        }
}

/* ---------------------------------------------------------------------- */
static void random_normal(const orc_params *p, double u1, double u2, double n[3])
{
    const double ct = 1.0 - u1 * (1.0 - cos(p->max_normal_angle));
    const double st = sqrt(fmax(0.0, 1.0 - ct * ct));
    const double ph = 2.0 * M_PI * u2;
    n[0] = st * cos(ph);
    n[1] = st * sin(ph);
    n[2] = ct;
}

static void perturb_normal(const orc_params *p, const double n[3], double amax, double u1,
                           double u2, double out[3])
{
    double h[3] = {1.0, 0.0, 0.0};
    if (fabs(n[0]) >= 0.9) { h[0] = 0.0; h[1] = 1.0; }
    double e1[3] = {n[1] * h[2] - n[2] * h[1], n[2] * h[0] - n[0] * h[2], n[0] * h[1] - n[1] * h[0]};
    const double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    e1[0] /= l1; e1[1] /= l1; e1[2] /= l1;
    const double e2[3] = {n[1] * e1[2] - n[2] * e1[1], n[2] * e1[0] - n[0] * e1[2],
                          n[0] * e1[1] - n[1] * e1[0]};
    const double al = amax * u1, ph = 2.0 * M_PI * u2;
    const double ca = cos(al), sa = sin(al), cp = cos(ph), sp = sin(ph);
    double m[3];
    for (int k = 0; k < 3; ++k) m[k] = ca * n[k] + sa * (cp * e1[k] + sp * e2[k]);
    const double lm = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    for (int k = 0; k < 3; ++k) m[k] /= lm;
    if (m[2] < cos(p->max_normal_angle)) { out[0] = n[0]; out[1] = n[1]; out[2] = n[2]; }
    else { out[0] = m[0]; out[1] = m[1]; out[2] = m[2]; }
}

void orc_init(const orc_view *v, const orc_params *p, int pass, double *depth, double *normal,
              double *cost, double *prob)
{
    const int W = v->width, H = v->height;
    const size_t HW = (size_t)W * H;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            double u[4];
            rng_block(v, p, (uint32_t)i, init_code(pass), 0u, u);
            depth[i] = p->depth_min + u[0] * (p->depth_max - p->depth_min);
            random_normal(p, u[1], u[2], normal + 3 * i);
            cost[i] = ORC_MAX_COST;
            for (int s = 0; s < v->n_src; ++s) prob[(size_t)s * HW + i] = 0.5;
        }
}

/* Reset view-selection state and costs at the start of a pass whose
 * hypotheses come from a previous pass. */
void orc_reset_state(const orc_view *v, double *cost, double *prob)
{
    const size_t HW = (size_t)v->width * v->height;
    for (size_t i = 0; i < HW; ++i) cost[i] = ORC_MAX_COST;
    for (size_t i = 0; i < HW * v->n_src; ++i) prob[i] = 0.5;
}

/* Counter word of a pixel's Monte-Carlo draws: the pixel index, or with
 * draw_group = g the index of its group of g x (128 / g) pixels (columns
 * g i .. g i + g - 1, rows (128/g) j .. (128/g) j + 128/g - 1), so every
 * pixel of a group uses the same uniforms (each against its own weights). */
static inline uint32_t draw_key(const orc_view *v, const orc_params *p, uint32_t pixel)
{
    if (!p->draw_group) return pixel;
    const uint32_t W = (uint32_t)v->width, gw = (uint32_t)p->draw_group, gh = 128u / gw;
    const uint32_t x = pixel % W, y = pixel / W;
    return (y / gh) * ((W + gw - 1u) / gw) + x / gw;
}

/* Monte-Carlo view selection: num_samples draws of a source with probability
 * q[s] / tot (inverse CDF over the sources in index order; a draw that
 * rounding leaves past the last boundary goes to the last source of positive
 * weight); cnt[s] = number of draws of s.  tot must be the sum of q in index
 * order.  With no positive weight every source counts once. */
static void draw_sources(const orc_view *v, const orc_params *p, uint32_t pixel, uint32_t code,
                         const double *q, double tot, int *cnt)
{
    const int S = v->n_src;
    for (int s = 0; s < S; ++s) cnt[s] = 0;
    if (!(tot > 0.0)) {
        for (int s = 0; s < S; ++s) cnt[s] = 1;
        return;
    }
    int last_pos = 0;
    for (int s = 0; s < S; ++s) if (q[s] > 0.0) last_pos = s;
    const uint32_t key = draw_key(v, p, pixel);
    for (int k = 0; k < p->num_samples; ++k) {
        double ub[4];
        rng_block(v, p, key, code, 2u + (uint32_t)(k >> 2), ub);
        const double target = ub[k & 3] * tot;
        double acc = 0.0;
        int pick = -1;
        for (int s = 0; s < S; ++s) {
            acc += q[s];
            if (target < acc) { pick = s; break; }
        }
        if (pick < 0) pick = last_pos;
        cnt[pick] += 1;
    }
}

/* The draw of one pixel for given weights (exported for the statistical
 * checks of the sampler). */
void orc_draw(const orc_view *v, const orc_params *p, uint32_t pixel, uint32_t code,
              const double *q, int32_t *cnt)
{
    double tot = 0.0;
    int c[ORC_MAX_SRC];
    for (int s = 0; s < v->n_src; ++s) tot += q[s];
    draw_sources(v, p, pixel, code, q, tot, c);
    for (int s = 0; s < v->n_src; ++s) cnt[s] = c[s];
}

void orc_random_normal(const orc_params *p, double u1, double u2, double n[3])
{
    random_normal(p, u1, u2, n);
}

void orc_perturb_normal(const orc_params *p, const double n[3], double amax, double u1, double u2,
                        double out[3])
{
    perturb_normal(p, n, amax, u1, u2, out);
}

static void propagate(const double *depth, const double *normal, int W, int xq, int yq, int x,
                      int y, double *d, double n[3])
{
    const size_t q = (size_t)yq * W + xq;
    const double *nq = normal + 3 * q;
    *d = depth[q] - (nq[0] * (x - xq) + nq[1] * (y - yq)) / nq[2];
    n[0] = nq[0]; n[1] = nq[1]; n[2] = nq[2];
}

/* One red-black half-step: update every pixel with (x+y)%2 == color.
 * winner (H*W ints) and cand (H*W*8 costs) are optional diagnostics. */
void orc_halfstep(const orc_view *v, const orc_params *p, int pass, int iter, int color,
                  double *depth, double *normal, double *cost, double *prob, int32_t *winner,
                  double *cand)
{
    const int W = v->width, H = v->height, S = v->n_src;
    const size_t HW = (size_t)W * H;
    const double decay = pow(p->perturb_decay, iter);
    const double range = p->depth_max - p->depth_min;
    const uint32_t code = step_code(pass, iter, color);
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < H; ++y) {
        for (int x = (y + color) & 1; x < W; x += 2) {
            const size_t i = (size_t)y * W + x;
            double hd[ORC_NHYP], hn[ORC_NHYP][3];
            hd[0] = depth[i];
            for (int k = 0; k < 3; ++k) hn[0][k] = normal[3 * i + k];
            /* adaptive checkerboard: best stored cost along each direction */
            static const int DX[4] = {-1, 1, 0, 0}, DY[4] = {0, 0, -1, 1};
            for (int dir = 0; dir < 4; ++dir) {
                int bx = -1, by = -1;
                double bc = 0.0;
                for (int k = 0; k < p->n_far; ++k) {
                    const int dist = 2 * k + 1;
                    const int qx = x + DX[dir] * dist, qy = y + DY[dir] * dist;
                    if (qx < 0 || qy < 0 || qx >= W || qy >= H) break;
                    const double c = cost[(size_t)qy * W + qx];
                    if (bx < 0 || c < bc) { bx = qx; by = qy; bc = c; }
                }
                if (bx < 0) {
                    hd[1 + dir] = hd[0];
                    for (int k = 0; k < 3; ++k) hn[1 + dir][k] = hn[0][k];
                } else {
                    propagate(depth, normal, W, bx, by, x, y, &hd[1 + dir], hn[1 + dir]);
                }
            }
            double u0[4], u1[4];
            rng_block(v, p, (uint32_t)i, code, 0u, u0);
            rng_block(v, p, (uint32_t)i, code, 1u, u1);
            hd[5] = p->depth_min + u0[0] * range;
            random_normal(p, u0[1], u0[2], hn[5]);
            hd[6] = clampd(hd[0] + (2.0 * u0[3] - 1.0) * p->perturbation * range * decay,
                           p->depth_min, p->depth_max);
            for (int k = 0; k < 3; ++k) hn[6][k] = hn[0][k];
            hd[7] = hd[0];
            perturb_normal(p, hn[0], p->normal_perturbation * decay, u1[0], u1[1], hn[7]);

            /* view selection from the current hypothesis */
            double c0[ORC_MAX_SRC], alpha[ORC_MAX_SRC], q[ORC_MAX_SRC];
            double jt[3];
            const double *J = v->J;
            for (int k = 0; k < 3; ++k)
                jt[k] = J[0 * 3 + k] * hn[0][0] + J[1 * 3 + k] * hn[0][1] + J[2 * 3 + k] * hn[0][2];
            const double jtn = sqrt(jt[0] * jt[0] + jt[1] * jt[1] + jt[2] * jt[2]);
            const double inv2n = 1.0 / (2.0 * p->ncc_sigma * p->ncc_sigma);
            const double inv2i = 1.0 / (2.0 * p->incident_angle_sigma * p->incident_angle_sigma);
            double tot = 0.0;
            for (int s = 0; s < S; ++s) {
                c0[s] = orc_photo_cost(v, p, x, y, hd[0], hn[0], s);
                double prev = 0.0;
                int np = 0;
                for (int dir = 0; dir < 4; ++dir) {
                    const int qx = x + DX[dir], qy = y + DY[dir];
                    if (qx < 0 || qy < 0 || qx >= W || qy >= H) continue;
                    prev += prob[(size_t)s * HW + (size_t)qy * W + qx];
                    ++np;
                }
                prev = np ? prev / np : 0.5;
                const double L = exp(-c0[s] * c0[s] * inv2n);
                const double a1 = L * (prev * 0.999 + (1.0 - prev) * 0.001);
                const double a0 = 0.5 * ((1.0 - prev) * 0.999 + prev * 0.001);
                alpha[s] = a1 / (a1 + a0);
                const double *gs = v->g + 3 * s;
                double ci = fabs(hn[0][0] * gs[0] + hn[0][1] * gs[1] + hn[0][2] * gs[2]) / jtn;
                ci = ci > 1.0 ? 1.0 : ci;
                const double inc = acos(ci);
                const double tm = v->tri[s] < p->min_tri_angle ? v->tri[s] : p->min_tri_angle;
                const double ptri = 1.0 - (tm - p->min_tri_angle) * (tm - p->min_tri_angle) /
                                              (p->min_tri_angle * p->min_tri_angle);
                q[s] = alpha[s] * ptri * exp(-inc * inc * inv2i);
                tot += q[s];
            }
            int cnt[ORC_MAX_SRC];
            draw_sources(v, p, (uint32_t)i, code, q, tot, cnt);
            double wsum = 0.0;
            for (int s = 0; s < S; ++s) wsum += cnt[s];
            double best_c = 0.0;
            int best = 0;
            for (int k = 0; k < ORC_NHYP; ++k) {
                double acc = 0.0;
                for (int s = 0; s < S; ++s) {
                    if (!cnt[s]) continue;
                    const double ph = k == 0 ? c0[s] : orc_photo_cost(v, p, x, y, hd[k], hn[k], s);
                    acc += cnt[s] * total_cost(v, p, pass, x, y, hd[k], hn[k], s, ph);
                }
                const double ck = acc / wsum;
                if (cand) cand[i * ORC_NHYP + k] = ck;
                if (k == 0 || ck < best_c) { best_c = ck; best = k; }
            }
            depth[i] = hd[best];
            for (int k = 0; k < 3; ++k) normal[3 * i + k] = hn[best][k];
            cost[i] = best_c;
            for (int s = 0; s < S; ++s) prob[(size_t)s * HW + i] = alpha[s];
            if (winner) winner[i] = best;
        }
    }
}

/* Filter: keep a pixel when at least min_num_consistent sources see it with
 * NCC >= min_ncc, triangulation angle >= min angle and (geometric pass)
 * forward-backward error <= filter_geom_max_cost; else depth 0 and normal 0.
 * mask (H*W uint64, optional) receives the consistent-source bits. */
void orc_filter(const orc_view *v, const orc_params *p, int pass, double *depth, double *normal,
                uint64_t *mask)
{
    const int W = v->width, H = v->height;
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            uint64_t m = 0;
            int count = 0;
            if (!(depth[i] > 0.0)) {
                if (mask) mask[i] = 0;
                continue;
            }
            for (int s = 0; s < v->n_src; ++s) {
                const double c = orc_photo_cost(v, p, x, y, depth[i], normal + 3 * i, s);
                if (!(1.0 - c >= p->filter_min_ncc)) continue;
                if (!(v->tri[s] >= p->filter_min_tri_angle)) continue;
                if (pass == 1 && !(orc_geo_error(v, p, x, y, depth[i], s) <= p->filter_geom_max_cost))
                    continue;
                m |= (uint64_t)1 << s;
                ++count;
            }
            if (mask) mask[i] = m;
            if (count < p->filter_min_num_consistent) {
                depth[i] = 0.0;
                normal[3 * i] = normal[3 * i + 1] = normal[3 * i + 2] = 0.0;
            }
        }
}

/* Full solve of one pass: init (pass 0) or reset (pass 1, hypotheses given),
 * num_iterations x {red, black}, optional filter. */
void orc_solve(const orc_view *v, const orc_params *p, int pass, int num_iterations, int init,
               int do_filter, double *depth, double *normal, double *cost, double *prob,
               uint64_t *mask)
{
    if (init) orc_init(v, p, pass, depth, normal, cost, prob);
    else orc_reset_state(v, cost, prob);
    for (int it = 0; it < num_iterations; ++it)
        for (int color = 0; color < 2; ++color)
            orc_halfstep(v, p, pass, it, color, depth, normal, cost, prob, NULL, NULL);
    if (do_filter) orc_filter(v, p, pass, depth, normal, mask);
}
