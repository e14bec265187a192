// vsm_pm.hip -- PatchMatch kernels for gfx950 (wave64).
//
// Work decomposition: one workgroup (256 threads, 4 waves) per 32x16 tile of
// one reference view; in a red-black half-step each thread owns one pixel of
// the active colour.  A launch covers the tiles of every view in a batch (the
// tile list maps blockIdx -> (view, tile)), so batched and one-at-a-time
// solves run the same per-pixel code and give bit-identical results.
//
// Numerics: per tile, the FP64 matrices M_s (ref (x,y,1,d) -> src) are
// re-centred on the tile centre and on an integer source anchor, then
// normalised and rounded to FP32; all per-tap arithmetic is FP32 on small
// coordinates (|x| ~ 1e2), which keeps source positions accurate to ~1e-5 px
// at 600 km camera range.  Compiled with -ffp-contract=off: every fused
// multiply-add is an explicit fmaf, so recomputed quantities are bit-identical.
//
// Throughput structure of a window pass (the ~80 % of the time spent in
// evaluating plane hypotheses): the whole wave reads one source image (the
// texture base is a scalar), the gathers of tap t+1 are issued before tap t
// is accumulated, and for windows of at most 36 taps the bilateral weights
// come from a per-pixel LDS table.  The kernel is pinned to 2 waves per SIMD.
//
// Semantics follow the oracle (oracle/pm_oracle.c): bilaterally weighted NCC,
// adaptive red-black propagation (best stored cost among 2k+1 steps in each of
// four directions), random and perturbed hypotheses, Monte-Carlo view
// selection with triangulation/incident-angle priors, geometric
// forward-backward term, filter.
#include <climits>

#include "vsm_device.h"
#include "vsm_kernels.h"

namespace vsm {

constexpr int kGroupH = 4;      // candidate hypotheses per shared window pass (4 + 3)
constexpr int kGroupV = 4;      // sources per shared window pass in view selection
constexpr int kFarChunk = 8;    // propagation candidates loaded together per direction
constexpr int kProbChunk = 4;   // sources whose neighbour probabilities are loaded together

// Largest relative change of the homogeneous depth q.z across a window for
// which the per-tap reciprocal is replaced by one Newton step from the exact
// centre reciprocal: iz = iz0 (2 - qz iz0), relative error delta^2 <= 1e-6.
// Satellite cameras (600 km range) give delta ~ 1e-5.
constexpr float kDeltaMax = 1e-3f;

// Depth range of the hypotheses a window pass may take without per-pixel
// bounds work: half the view's depth range beyond either end (propagated
// planes extrapolate past [depth_min, depth_max]).
__device__ __forceinline__ float hyp_depth_lo(const ViewDesc &V)
{
    return V.depth_min - 0.5f * (V.depth_max - V.depth_min);
}
__device__ __forceinline__ float hyp_depth_hi(const ViewDesc &V)
{
    return V.depth_max + 0.5f * (V.depth_max - V.depth_min);
}

struct SrcLds {
    float A[12];                 // tile-centred ref -> src (anchor-relative), normalised
    float B[12];                 // src (anchor-relative) -> ref (tile-centre-relative)
    int ax, ay, W, H;
    int pitch;                   // W + 1: texels per row of the pair texture
    int base;                    // ay * pitch + ax: texel offset of the anchor
    float lox, hix, loy, hiy;    // in-image box in anchor-relative coordinates
    const uint8_t *pair;         // pair texture (k_pack_pair)
    const float4 *depth4;        // photometric depth of the source, 2x2 per texel
    float ptri, tri, g0, g1, g2;
    int inner;                   // every window of the tile lies inside the source (see below)
};

struct Proj {
    float q0x, q0y, q0z;         // centre projection
    float hux, huy, huz;         // d q / d dx
    float hvx, hvy, hvz;         // d q / d dy
};

// Hypotheses a window pass may take without per-pixel bounds work: depth
// within half the depth range outside [depth_min, depth_max], plane slopes
// |a|, |b| at most those of the steepest normal the search generates.
__device__ __forceinline__ float slope_bound(const PMConst &c)
{
    const float cm = c.cos_max_normal;
    return cm > 0.f ? 1.01f * sqrtf(fmaxf(1.f - cm * cm, 0.f)) / cm : 0.f;
}

// 1 when every tap of every window of a tile (pixel box hu x hv around the
// tile centre, plus the window radius) under every hypothesis with depth in
// [dlo, dhi] and slopes |a|, |b| <= amax lands inside the source with a
// 1.5-px guard, and the linearised per-tap reciprocal is accurate there
// (delta <= kDeltaMax / 2).  The taps (u, v, 1, d + a dx + b dy) fill the box
// |u| <= hu + r, |v| <= hv + r, d in [dlo - 2 amax r, dhi + 2 amax r]; with a
// positive homogeneous depth at its 8 corners (it is affine there) their
// projections lie in the convex hull of the corners' projections.
__device__ int source_inner(const double (&A)[12], double hu, double hv, double r, double dlo,
                            double dhi, double amax, double lox, double hix, double loy, double hiy)
{
    const double d0 = dlo - 2.0 * amax * r, d1 = dhi + 2.0 * amax * r;
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300, zmin = 1e300;
    for (int k = 0; k < 8; ++k) {
        const double u = (k & 1) ? hu + r : -(hu + r);
        const double v = (k & 2) ? hv + r : -(hv + r);
        const double d = (k & 4) ? d1 : d0;
        const double qz = A[8] * u + A[9] * v + A[10] + A[11] * d;
        if (!(qz > 0.0)) return 0;
        const double px = (A[0] * u + A[1] * v + A[2] + A[3] * d) / qz;
        const double py = (A[4] * u + A[5] * v + A[6] + A[7] * d) / qz;
        xmin = fmin(xmin, px); xmax = fmax(xmax, px);
        ymin = fmin(ymin, py); ymax = fmax(ymax, py);
        zmin = fmin(zmin, qz);
    }
    const double g = 1.5;
    if (!(xmin >= lox + g && xmax <= hix - g && ymin >= loy + g && ymax <= hiy - g)) return 0;
    const double delta = r * (fabs(A[8]) + fabs(A[9]) + 2.0 * amax * fabs(A[11])) / zmin;
    return delta <= 0.5 * (double)kDeltaMax ? 1 : 0;
}

// FP64 tile re-centring of one source's matrices (run by one thread per source).
// hu, hv: largest pixel offset from the tile centre.
__device__ void setup_source(const ViewDesc &V, const PMConst &c, int s, int uc, int vc, int hu, int hv,
                             SrcLds &o)
{
    const double *M = V.M[s];
    const double *Mb = V.Mb[s];
    const double dmid = 0.5 * ((double)V.depth_min + (double)V.depth_max);
    const double qx = M[0] * uc + M[1] * vc + M[2] + M[3] * dmid;
    const double qy = M[4] * uc + M[5] * vc + M[6] + M[7] * dmid;
    const double qz = M[8] * uc + M[9] * vc + M[10] + M[11] * dmid;
    int ax, ay;
    source_anchor(qx, qy, qz, V.src_w[s], V.src_h[s], ax, ay);
    double A[12], B[12];
    for (int r = 0; r < 3; ++r) {
        A[4 * r + 0] = M[4 * r + 0];
        A[4 * r + 1] = M[4 * r + 1];
        A[4 * r + 2] = M[4 * r + 0] * uc + M[4 * r + 1] * vc + M[4 * r + 2];
        A[4 * r + 3] = M[4 * r + 3];
        B[4 * r + 0] = Mb[4 * r + 0];
        B[4 * r + 1] = Mb[4 * r + 1];
        B[4 * r + 2] = Mb[4 * r + 0] * ax + Mb[4 * r + 1] * ay + Mb[4 * r + 2];
        B[4 * r + 3] = Mb[4 * r + 3];
    }
    for (int k = 0; k < 4; ++k) {
        A[k] -= ax * A[8 + k];
        A[4 + k] -= ay * A[8 + k];
        B[k] -= uc * B[8 + k];
        B[4 + k] -= vc * B[8 + k];
    }
    double ma = 0.0, mb = 0.0;
    for (int k = 0; k < 12; ++k) { ma = fmax(ma, fabs(A[k])); mb = fmax(mb, fabs(B[k])); }
    ma = ma > 0.0 ? 1.0 / ma : 1.0;
    mb = mb > 0.0 ? 1.0 / mb : 1.0;
    for (int k = 0; k < 12; ++k) { o.A[k] = (float)(A[k] * ma); o.B[k] = (float)(B[k] * mb); }
    o.ax = ax; o.ay = ay;
    o.W = V.src_w[s]; o.H = V.src_h[s];
    o.pitch = o.W + 1;
    o.base = ay * o.pitch + ax;
    o.lox = (float)(-ax); o.hix = (float)(o.W - 1 - ax);
    o.loy = (float)(-ay); o.hiy = (float)(o.H - 1 - ay);
    o.pair = V.src_pair[s];
    o.depth4 = V.src_depth4[s];
    o.ptri = V.ptri[s]; o.tri = V.tri[s];
    o.g0 = V.g[s][0]; o.g1 = V.g[s][1]; o.g2 = V.g[s][2];
    {
        const double dr = (double)V.depth_max - (double)V.depth_min;
        const double dlo = hyp_depth_lo(V) - 1e-3 * (fabs(dr) + 1.0);
        const double dhi = hyp_depth_hi(V) + 1e-3 * (fabs(dr) + 1.0);
        o.inner = source_inner(A, hu, hv, (double)c.radius, dlo, dhi, (double)slope_bound(c),
                               (double)o.lox, (double)o.hix, (double)o.loy, (double)o.hiy);
    }
}

__device__ __forceinline__ Proj make_proj(const SrcLds &S, float du, float dv, float d, float a,
                                          float b)
{
    Proj P;
    P.q0x = fmaf(S.A[0], du, fmaf(S.A[1], dv, fmaf(S.A[3], d, S.A[2])));
    P.q0y = fmaf(S.A[4], du, fmaf(S.A[5], dv, fmaf(S.A[7], d, S.A[6])));
    P.q0z = fmaf(S.A[8], du, fmaf(S.A[9], dv, fmaf(S.A[11], d, S.A[10])));
    P.hux = fmaf(a, S.A[3], S.A[0]);
    P.huy = fmaf(a, S.A[7], S.A[4]);
    P.huz = fmaf(a, S.A[11], S.A[8]);
    P.hvx = fmaf(b, S.A[3], S.A[1]);
    P.hvy = fmaf(b, S.A[7], S.A[5]);
    P.hvz = fmaf(b, S.A[11], S.A[9]);
    return P;
}

typedef const __attribute__((address_space(1))) uint8_t *gbyte_t;
typedef const __attribute__((address_space(1))) float *gfloat_t;
typedef const __attribute__((address_space(3))) uint32_t *lquad_t;
// A 4-byte word at a 2-byte boundary of the pair texture: the HSA runtime
// enables unaligned global access, so this is one global_load_dword.
typedef uint32_t u32_any __attribute__((aligned(1)));
typedef const __attribute__((address_space(1))) u32_any *gword_t;

// floor(x) as an integer in one instruction (x finite).
__device__ __forceinline__ int cvt_floor(float x)
{
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// The fields of a source the window pass reads per tap (copied out of LDS
// once per evaluation).
struct SrcView {
    float lox, hix, loy, hiy;
    int pitch, base;             // in texels
    int pitchb, baseb;           // row pitch and anchor offset in bytes
    const uint8_t *pair;
};

// Every lane of a wave reads the same source in a window pass, so the texture
// pointer, row pitch and anchor offset go to scalar registers and each gather
// is a scalar base plus a 32-bit byte offset (2 * texel index < 2^32 for
// images of at most 2^30 pixels).
__device__ __forceinline__ SrcView view_of(const SrcLds &S)
{
    SrcView v;
    v.lox = S.lox; v.hix = S.hix; v.loy = S.loy; v.hiy = S.hiy;
    v.pitch = S.pitch; v.base = S.base;
    v.pitchb = __builtin_amdgcn_readfirstlane(S.pitch * 2);
    v.baseb = __builtin_amdgcn_readfirstlane((int)(2u * (uint32_t)S.base));   // may exceed 2^31
    const uint64_t a = (uint64_t)S.pair;
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(a >> 32));
    v.pair = (const uint8_t *)(lo | (hi << 32));
    return v;
}

// Source samples are blended in units of 2^-24 grey levels (blend_quad).
constexpr float kQuadScale = 0x1p-24f;

// Image value of the texel containing anchor-relative (px, py) (inside the
// image), times kQuadScale: the offset the source samples of a window are
// centred on.
__device__ __forceinline__ float texel_at(const SrcView &S, float px, float py)
{
    const int idx = __mul24(cvt_floor(py), S.pitch) + (cvt_floor(px) + S.base);
    return (float)((gbyte_t)S.pair)[2u * (unsigned)idx] * kQuadScale;
}

// Bilateral weight of a tap: 2^-(r^2 ccc + (dx^2+dy^2) css), the hardware
// exp2 (weights below 2^-126 flush to zero, which no sum can notice).
__device__ __forceinline__ float tap_weight(const PMConst &c, float r, int dx, int dy)
{
    return __builtin_amdgcn_exp2f(-fmaf(r * r, c.ccc, (float)(dx * dx + dy * dy) * c.css));
}

struct RefStats {
    float rc, sw, inv_sw, mr, var_r;
    const float *wt;             // tap weights in LDS (stride kThreads), or null
    int ldsw;                    // row pitch of the reference tile in LDS
};

// Windows of at most kWTab taps keep each pixel's tap weights in LDS (written
// once per half-step by ref_stats, read by every window pass of the pixel;
// 36 KB per workgroup for 6 x 6 taps).
constexpr int kWTab = 36;

__device__ __forceinline__ RefStats ref_stats(const PMConst &c, const float *s_ref, int lbase, int ldsw,
                                              float *wt = nullptr)
{
    RefStats st;
    st.rc = s_ref[lbase];
    st.wt = wt;
    st.ldsw = ldsw;
    float sw = 0.f, sr = 0.f, srr = 0.f;
    int t = 0;
    for (int j = 0; j < c.ntaps; ++j) {
        const int dy = -c.radius + j * c.step;
        const float *row = s_ref + lbase + dy * ldsw;
        for (int i = 0; i < c.ntaps; ++i, ++t) {
            const int dx = -c.radius + i * c.step;
            const float r = row[dx] - st.rc;
            const float w = tap_weight(c, r, dx, dy);
            if (wt) wt[t * kThreads] = w;
            sw += w;
            sr = fmaf(w, r, sr);
            srr = fmaf(w * r, r, srr);
        }
    }
    st.sw = sw;
    st.inv_sw = 1.f / sw;
    st.mr = sr * st.inv_sw;
    st.var_r = srr * st.inv_sw - st.mr * st.mr;
    return st;
}

// One bilinear tap split in two: tap_issue forms the anchor-relative source
// position of column offset fi and starts the gather of its quad, the word
// of texels (ix, iy) and (ix + 1, iy) of the pair texture; blend_quad
// interpolates it (DN).  The byte offset 2 (iy*pitch + ix) is formed with a
// signed 24-bit multiply-add on anchor-relative integers plus the anchor's
// offset 2 (ay*pitch + ax) (|relative coordinates| < 2^23, image side <=
// 32768).  With LDSRC the quad comes from a box of the source staged in LDS
// as aligned words (S.pitchb and S.baseb are then the box pitch and the
// anchor's index in it, in words).
template <bool EXACT, bool CLAMP, bool LDSRC = false>
__device__ __forceinline__ void tap_issue(const SrcView &S, const Proj &P, float giu, float rx, float ry,
                                          float rz, float fi, uint32_t &v, float &fx, float &fy)
{
    const float qx = fmaf(fi, P.hux, rx);
    const float qy = fmaf(fi, P.huy, ry);
    const float iz = EXACT ? __builtin_amdgcn_rcpf(fmaf(fi, P.huz, rz)) : fmaf(fi, giu, rz);
    const float px = CLAMP ? __builtin_amdgcn_fmed3f(qx * iz, S.lox, S.hix) : qx * iz;
    const float py = CLAMP ? __builtin_amdgcn_fmed3f(qy * iz, S.loy, S.hiy) : qy * iz;
    const int ix = cvt_floor(px), iy = cvt_floor(py);
    fx = __builtin_amdgcn_fractf(px);
    fy = __builtin_amdgcn_fractf(py);
    if constexpr (LDSRC) {
        v = ((lquad_t)S.pair)[__mul24(iy, S.pitchb) + ix + S.baseb];
    } else {
        // modulo-2^32 arithmetic; the sum is the byte offset, below 2^32
        const unsigned off = (unsigned)__mul24(iy, S.pitchb) + (((unsigned)ix << 1) + (unsigned)S.baseb);
        v = *(gword_t)((gbyte_t)S.pair + off);
    }
}

// Bilinear value of the quad (bytes a = I(x,y), c = I(x,y+1), b = I(x+1,y),
// d = I(x+1,y+1)), times kQuadScale:
// top = fma(fx, b - a, a), bot = fma(fx, d - c, c), fma(fy, bot - top, top).
// The bytes go into the low bits of binary16 lanes, where they read as the
// subnormals a 2^-24 .. d 2^-24 (exact); one packed binary16 subtraction
// forms (b - a, d - c) 2^-24 (exact), and the row FMAs take their binary16
// operands as mixed-precision sources.  Each FP32 operation sees the
// unscaled blend's operands times 2^-24 and no result falls below the
// smallest normal float, so the value is the unscaled blend's times 2^-24
// exactly (2 to 3 fewer VALU instructions per tap than converting the four
// bytes).  The window's source moment sums then carry kQuadScale and
// kQuadScale^2, which the NCC cancels.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float blend_quad(uint32_t v, float fx, float fy)
{
    const half2_t ac = __builtin_bit_cast(half2_t, __builtin_amdgcn_perm(0u, v, 0x0c010c00u)); // a | c << 16
    const half2_t bd = __builtin_bit_cast(half2_t, __builtin_amdgcn_perm(0u, v, 0x0c030c02u)); // b | d << 16
    const half2_t df = bd - ac;
    const float top = fmaf(fx, (float)df.x, (float)ac.x);
    const float bot = fmaf(fx, (float)df.y, (float)ac.y);
    return fmaf(fy, bot - top, top);
}

// Row start of row offset dy for each hypothesis: q (or, without EXACT, the
// linearised 1/q.z) at column offset -r.
template <int G, bool EXACT>
__device__ __forceinline__ void row_start(const Proj (&P)[G], const float (&iz0)[G],
                                          const float (&giu)[G], const float (&giv)[G], float frad,
                                          int dy, float (&rx)[G], float (&ry)[G], float (&rz)[G])
{
    const float fdy = (float)dy;
#pragma unroll
    for (int k = 0; k < G; ++k) {
        rx[k] = fmaf(fdy, P[k].hvx, fmaf(frad, P[k].hux, P[k].q0x));
        ry[k] = fmaf(fdy, P[k].hvy, fmaf(frad, P[k].huy, P[k].q0y));
        rz[k] = EXACT ? fmaf(fdy, P[k].hvz, fmaf(frad, P[k].huz, P[k].q0z))
                      : fmaf(fdy, giv[k], fmaf(frad, giu[k], iz0[k]));
    }
}

// Window pass of G plane hypotheses against the same source S: the bilateral
// weight and the reference value of each tap are computed once and shared;
// the G gathers per tap are independent.  EXACT selects the per-tap
// reciprocal (fallback when some lane's window exceeds kDeltaMax).
template <int G, bool EXACT, bool CLAMP, bool WT, bool LDSRC = false>
__device__ __forceinline__ void window_pass(const PMConst &c, const SrcView (&SV)[G],
                                            const Proj (&P)[G], const float (&iz0)[G],
                                            const float (&sc)[G], const float *s_ref, int lbase,
                                            const RefStats &R, float (&ss)[G], float (&sss)[G],
                                            float (&srs)[G])
{
    const int rad = c.radius, st = c.step, nt = c.ntaps;
    const float frad = (float)(-rad);
    float giu[G], giv[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        giu[k] = -iz0[k] * iz0[k] * P[k].huz;
        giv[k] = -iz0[k] * iz0[k] * P[k].hvz;
        ss[k] = 0.f; sss[k] = 0.f; srs[k] = 0.f;
    }
    // Taps in row-major order as one sequence; the gathers of tap t+1
    // are issued before tap t is accumulated, with two register sets
    // used in turn.  The issue past the last tap repeats the last tap.
    const int T = nt * nt;
    int ii = 0, jj = 0;                       // next tap to issue
    float rx[G], ry[G], rz[G];
    row_start<G, EXACT>(P, iz0, giu, giv, frad, -rad, rx, ry, rz);
    uint32_t va[G], vb[G];
    float xa[G], ya[G], xb[G], yb[G];
    // tok: values the gathers must wait for (empty asm statements tie
    // them to the tap position, so the compiler keeps the issue after
    // the accumulation of the set being reloaded)
    // The tap's reference value and bilateral weight are read from LDS with
    // its gathers, one accumulation ahead of their use.
    auto issue = [&](uint32_t (&v)[G], float (&fx)[G], float (&fy)[G], const float (&tok)[G], float &rr,
                     float &ww) {
        float fi = (float)(ii * st);
        int la = lbase + (-rad + jj * st) * R.ldsw + (-rad + ii * st);
        int wa = (jj * nt + ii) * kThreads;
#pragma unroll
        for (int k = 0; k < G; ++k) asm volatile("" : "+v"(fi), "+v"(la), "+v"(wa) : "v"(tok[k]));
        rr = s_ref[la];
        if constexpr (WT) ww = R.wt[wa];
        // a wave issuing gathers goes ahead of one accumulating, so its
        // loads leave early (-5 % per stride-1 half-step, neutral at stride 2)
        __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int k = 0; k < G; ++k)
            tap_issue<EXACT, CLAMP, LDSRC>(SV[k], P[k], giu[k], rx[k], ry[k], rz[k], fi, v[k], fx[k], fy[k]);
        __builtin_amdgcn_s_setprio(0);
        if (ii + 1 < nt) {
            ++ii;
        } else if (jj + 1 < nt) {
            ii = 0;
            ++jj;
            row_start<G, EXACT>(P, iz0, giu, giv, frad, -rad + jj * st, rx, ry, rz);
        }
    };
    int ic = 0, jc = 0;                       // next tap to accumulate
    auto consume = [&](const uint32_t (&v)[G], const float (&fx)[G], const float (&fy)[G], float rr, float ww) {
        const int dx = -rad + ic * st, dy = -rad + jc * st;
        const float r = rr - R.rc;
        const float w = WT ? ww : tap_weight(c, r, dx, dy);
        const float wr = w * r;
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const float sv = blend_quad(v[k], fx[k], fy[k]) - sc[k];
            const float ws = w * sv;
            ss[k] += ws;
            sss[k] = fmaf(ws, sv, sss[k]);
            srs[k] = fmaf(wr, sv, srs[k]);
        }
        if (++ic == nt) { ic = 0; ++jc; }
    };
    float ra = 0.f, wa = 0.f, rb = 0.f, wb = 0.f;
    issue(va, xa, ya, ss, ra, wa);
    int t = 0;
#pragma unroll 1
    for (; t + 1 < T; t += 2) {
        issue(vb, xb, yb, ss, rb, wb);
        consume(va, xa, ya, ra, wa);
        // set a's next gathers come after its last use (no register copy
        // at the loop edge, which would wait for them)
        issue(va, xa, ya, ss, ra, wa);
        consume(vb, xb, yb, rb, wb);
    }
    if (t < T) consume(va, xa, ya, ra, wa);
}

// window_pass for the 6 x 6-tap window (radius 5, step 2) with tap weights in
// LDS: the taps of a row are unrolled, so tap offsets, LDS offsets and the
// register set of each tap are constants; the rows run in a loop.  Same
// arithmetic in the same order as window_pass, so the same sums.
template <int G, bool CLAMP>
__device__ __forceinline__ void window_pass_r5s2(const PMConst &c, const SrcView (&SV)[G],
                                                 const Proj (&P)[G], const float (&iz0)[G],
                                                 const float (&sc)[G], const float *s_ref, int lbase,
                                                 const RefStats &R, float (&ss)[G], float (&sss)[G],
                                                 float (&srs)[G])
{
    constexpr int NT = 6, ST = 2, RAD = 5;
    float giu[G], giv[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        giu[k] = -iz0[k] * iz0[k] * P[k].huz;
        giv[k] = -iz0[k] * iz0[k] * P[k].hvz;
        ss[k] = 0.f; sss[k] = 0.f; srs[k] = 0.f;
    }
    float rx[G], ry[G], rz[G];
    row_start<G, false>(P, iz0, giu, giv, (float)(-RAD), -RAD, rx, ry, rz);
    uint32_t v[2][G];
    float fx[2][G], fy[2][G], rr[2], ww[2];
    const float *ref_row = s_ref + lbase - RAD * R.ldsw - RAD;    // row -RAD, column -RAD
    const float *wt_row = R.wt;
    // gathers and LDS reads of tap ii of the current row into set ii & 1,
    // after the accumulation that last used that set (tok)
    auto issue = [&](int ii, const float (&tok)[G]) {
        const int b = ii & 1;
        float fi = (float)(ii * ST);
#pragma unroll
        for (int k = 0; k < G; ++k) asm volatile("" : "+v"(fi) : "v"(tok[k]));
        rr[b] = ref_row[ii * ST];
        ww[b] = wt_row[ii * kThreads];
        __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int k = 0; k < G; ++k)
            tap_issue<false, CLAMP>(SV[k], P[k], giu[k], rx[k], ry[k], rz[k], fi, v[b][k], fx[b][k], fy[b][k]);
        __builtin_amdgcn_s_setprio(0);
    };
    auto consume = [&](int ii) {
        const int b = ii & 1;
        const float r = rr[b] - R.rc;
        const float w = ww[b];
        const float wr = w * r;
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const float sv = blend_quad(v[b][k], fx[b][k], fy[b][k]) - sc[k];
            const float ws = w * sv;
            ss[k] += ws;
            sss[k] = fmaf(ws, sv, sss[k]);
            srs[k] = fmaf(wr, sv, srs[k]);
        }
    };
    issue(0, ss);
#pragma unroll 1
    for (int jj = 0; jj < NT; ++jj) {
#pragma unroll
        for (int ii = 0; ii < NT; ++ii) {
            if (ii + 1 < NT) {
                issue(ii + 1, ss);
            } else if (jj + 1 < NT) {
                // first tap of the next row into set 0 (free: tap 4 was accumulated)
                row_start<G, false>(P, iz0, giu, giv, (float)(-RAD), -RAD + (jj + 1) * ST, rx, ry, rz);
                ref_row += ST * R.ldsw;
                wt_row += NT * kThreads;
                issue(0, ss);
            }
            consume(ii);
        }
    }
}

// NCC cost of G plane hypotheses.  Each hypothesis' source samples are
// centred on the source texel at its window centre, the reference samples on
// the reference centre value (NCC is invariant to the offsets; they keep the
// FP32 moment sums well conditioned when the two views differ in brightness).
// SS[k] is the source of hypothesis k: all the same source (several planes
// against one source) or G sources of one plane (view selection).
// With LDSRC the window taps read the LDS box `staged` describes (its quad
// is the LDS array; see tap_issue) instead of the source in global memory;
// the centre texel still comes from global memory.  Same arithmetic, same
// result.
// lane_ok: this lane's hypotheses have depths in [hyp_depth_lo, hyp_depth_hi]
// and slopes within slope_bound.  When every lane's are and every source is
// `inner` for the tile, each window is known to lie inside its source with an
// accurate linearised reciprocal, and the per-lane bounds work is skipped: the
// pass then takes the variant the per-lane tests would (or the clamped one,
// whose clamp changes no interior coordinate), so the costs are the same.
template <int G, bool LDSRC = false>
__device__ __forceinline__ void ncc_cost_group(const PMConst &c, const SrcLds *const (&SS)[G],
                                               const Proj (&P)[G], const float *s_ref, int lbase,
                                               const RefStats &R, float (&cost)[G],
                                               const SrcView *staged = nullptr, bool lane_ok = false)
{
    bool valid[G];
    float iz0[G], sc[G], ss[G], sss[G], srs[G];
    SrcView SV[G];
    bool far = false, edge = false;
    const float fr = (float)c.radius;
    bool inner = !LDSRC;
#pragma unroll
    for (int k = 0; k < G; ++k) inner = inner && SS[k]->inner != 0;
    if (inner && __ballot(!lane_ok) == 0ull) {
#pragma unroll
        for (int k = 0; k < G; ++k) {
            SV[k] = view_of(*SS[k]);
            valid[k] = true;
            iz0[k] = __builtin_amdgcn_rcpf(P[k].q0z);
            sc[k] = texel_at(SV[k], P[k].q0x * iz0[k], P[k].q0y * iz0[k]);
        }
    } else {
#pragma unroll
        for (int k = 0; k < G; ++k) {
            SV[k] = view_of(*SS[k]);
            const SrcView &S = SV[k];
            valid[k] = P[k].q0z > 0.f;
            iz0[k] = valid[k] ? __builtin_amdgcn_rcpf(P[k].q0z) : 1.f;
            const float cx = P[k].q0x * iz0[k], cy = P[k].q0y * iz0[k];
            valid[k] = valid[k] && cx >= S.lox && cx <= S.hix && cy >= S.loy && cy <= S.hiy;
            sc[k] = valid[k] ? texel_at(S, cx, cy) : 0.f;
            const float delta = fr * (fabsf(P[k].huz) + fabsf(P[k].hvz)) * fabsf(iz0[k]);
            far = far || !(delta <= kDeltaMax);
            // conservative half-extent of the projected window (+1 px guard)
            const float jxx = (P[k].hux - cx * P[k].huz) * iz0[k], jxy = (P[k].hvx - cx * P[k].hvz) * iz0[k];
            const float jyx = (P[k].huy - cy * P[k].huz) * iz0[k], jyy = (P[k].hvy - cy * P[k].hvz) * iz0[k];
            const float mx = fr * (fabsf(jxx) + fabsf(jxy)) * (1.f + 2.f * delta) + 1.f;
            const float my = fr * (fabsf(jyx) + fabsf(jyy)) * (1.f + 2.f * delta) + 1.f;
            edge = edge || !(cx - mx >= S.lox && cx + mx <= S.hix && cy - my >= S.loy && cy + my <= S.hiy);
        }
    }
    if constexpr (LDSRC) {
        static_assert(G == 1, "one staged source per pass");
        const SrcView LV[1] = {*staged};
        if (__ballot(far) != 0ull)
            window_pass<1, true, true, false, true>(c, LV, P, iz0, sc, s_ref, lbase, R, ss, sss, srs);
        else if (__ballot(edge) != 0ull)
            window_pass<1, false, true, false, true>(c, LV, P, iz0, sc, s_ref, lbase, R, ss, sss, srs);
        else
            window_pass<1, false, false, false, true>(c, LV, P, iz0, sc, s_ref, lbase, R, ss, sss, srs);
    } else if (__ballot(far) != 0ull) {
        window_pass<G, true, true, false>(c, SV, P, iz0, sc, s_ref, lbase, R, ss, sss, srs);
    } else if (R.wt != nullptr && c.radius == 5 && c.step == 2) {
        if (__ballot(edge) != 0ull)
            window_pass_r5s2<G, true>(c, SV, P, iz0, sc, s_ref, lbase, R, ss, sss, srs);
        else
            window_pass_r5s2<G, false>(c, SV, P, iz0, sc, s_ref, lbase, R, ss, sss, srs);
    } else if (R.wt != nullptr) {
        if (__ballot(edge) != 0ull)
            window_pass<G, false, true, true>(c, SV, P, iz0, sc, s_ref, lbase, R, ss, sss, srs);
        else
            window_pass<G, false, false, true>(c, SV, P, iz0, sc, s_ref, lbase, R, ss, sss, srs);
    } else {
        if (__ballot(edge) != 0ull)
            window_pass<G, false, true, false>(c, SV, P, iz0, sc, s_ref, lbase, R, ss, sss, srs);
        else
            window_pass<G, false, false, false>(c, SV, P, iz0, sc, s_ref, lbase, R, ss, sss, srs);
    }
#pragma unroll
    for (int k = 0; k < G; ++k) {
        const float ms = ss[k] * R.inv_sw;
        const float var_s = sss[k] * R.inv_sw - ms * ms;
        const float cov = srs[k] * R.inv_sw - R.mr * ms;
        // var_s is in units of kQuadScale^2 (NCC itself is scale-free)
        const bool flat = R.var_r < c.min_var || var_s < c.min_var * (kQuadScale * kQuadScale);
        const float ncc = cov * __builtin_amdgcn_rsqf(R.var_r * var_s);
        cost[k] = (!valid[k] || flat) ? kMaxCost : fminf(fmaxf(1.f - ncc, 0.f), kMaxCost);
    }
}

// 1 - bilaterally weighted NCC of the plane P against source S.
__device__ __forceinline__ float ncc_cost(const PMConst &c, const SrcLds &S, const Proj &P,
                                          const float *s_ref, int lbase, const RefStats &R)
{
    const Proj PP[1] = {P};
    const SrcLds *const SS[1] = {&S};
    float cost[1];
    ncc_cost_group<1>(c, SS, PP, s_ref, lbase, R, cost);
    return cost[0];
}

// Forward-backward reprojection error (px) through the source depth map.
__device__ float geo_error(const PMConst &c, const SrcLds &S, const Proj &P, float du, float dv)
{
    const float gmax = c.geom_max;
    if (S.depth4 == nullptr) return gmax;
    if (!(P.q0z > 0.f)) return gmax;
    const float iq = __builtin_amdgcn_rcpf(P.q0z);
    const float px = P.q0x * iq, py = P.q0y * iq;
    if (!(px >= S.lox && px <= S.hix && py >= S.loy && py <= S.hiy)) return gmax;
    const float fx0 = floorf(px), fy0 = floorf(py);
    const int x0 = S.ax + (int)fx0, y0 = S.ay + (int)fy0;
    const float fx = px - fx0, fy = py - fy0;
    // one 16-byte load: d(x0,y0), d(x1,y0), d(x0,y1), d(x1,y1) with the
    // neighbours clamped to the image
    const float4 q = S.depth4[y0 * S.W + x0];
    const float da = q.x, db = q.y, dc = q.z, de = q.w;
    if (!(da > 0.f && db > 0.f && dc > 0.f && de > 0.f)) return gmax;
    const float top = fmaf(fx, db - da, da), bot = fmaf(fx, de - dc, dc);
    const float ds = fmaf(fy, bot - top, top);
    const float bx = fmaf(S.B[0], px, fmaf(S.B[1], py, fmaf(S.B[3], ds, S.B[2])));
    const float by = fmaf(S.B[4], px, fmaf(S.B[5], py, fmaf(S.B[7], ds, S.B[6])));
    const float bz = fmaf(S.B[8], px, fmaf(S.B[9], py, fmaf(S.B[11], ds, S.B[10])));
    if (!(bz > 0.f)) return gmax;
    const float ib = __builtin_amdgcn_rcpf(bz);
    const float ex = bx * ib - du, ey = by * ib - dv;
    return sqrtf(ex * ex + ey * ey);
}

__device__ __forceinline__ float total_cost(const PMConst &c, const SrcLds &S, const Proj &P,
                                            float photo, int pass, float du, float dv)
{
    if (pass == 1) {
        const float e = geo_error(c, S, P, du, dv);
        return photo + c.geom_reg * fminf(e, c.geom_max);
    }
    return photo;
}

__device__ __forceinline__ void random_normal(const PMConst &c, float u1, float u2, float n[3])
{
    // sin from (1 - cos)(1 + cos): no cancellation near the cone axis
    const float omc = u1 * c.one_minus_cos_max;
    const float ct = 1.f - omc;
    const float st = sqrtf(omc * (2.f - omc));
    const float ph = 6.283185307179586f * u2;
    float sp, cp;
    sincosf(ph, &sp, &cp);
    n[0] = st * cp;
    n[1] = st * sp;
    n[2] = ct;
}

__device__ __forceinline__ void perturb_normal(const PMConst &c, const float n[3], float amax,
                                               float u1, float u2, float out[3])
{
    float h0 = 1.f, h1 = 0.f;
    if (fabsf(n[0]) >= 0.9f) { h0 = 0.f; h1 = 1.f; }
    // e1 = n x h, h = (h0, h1, 0)
    float e1x = n[1] * 0.f - n[2] * h1, e1y = n[2] * h0 - n[0] * 0.f, e1z = n[0] * h1 - n[1] * h0;
    const float l1 = sqrtf(e1x * e1x + e1y * e1y + e1z * e1z);
    e1x /= l1; e1y /= l1; e1z /= l1;
    const float e2x = n[1] * e1z - n[2] * e1y, e2y = n[2] * e1x - n[0] * e1z,
                e2z = n[0] * e1y - n[1] * e1x;
    const float al = amax * u1, ph = 6.283185307179586f * u2;
    float sa, ca, sp, cp;
    sincosf(al, &sa, &ca);
    sincosf(ph, &sp, &cp);
    float m0 = ca * n[0] + sa * (cp * e1x + sp * e2x);
    float m1 = ca * n[1] + sa * (cp * e1y + sp * e2y);
    float m2 = ca * n[2] + sa * (cp * e1z + sp * e2z);
    const float lm = sqrtf(m0 * m0 + m1 * m1 + m2 * m2);
    m0 /= lm; m1 /= lm; m2 /= lm;
    if (m2 < c.cos_max_normal) { out[0] = n[0]; out[1] = n[1]; out[2] = n[2]; }
    else { out[0] = m0; out[1] = m1; out[2] = m2; }
}

// Sampling weight of source S for the current hypothesis (alpha * priors).
__device__ __forceinline__ float view_weight(const PMConst &c, const SrcLds &S, float alpha,
                                             const float n[3], float jtn)
{
    float ci = fabsf(n[0] * S.g0 + n[1] * S.g1 + n[2] * S.g2) / jtn;
    ci = fminf(ci, 1.f);
    const float inc = acosf(ci);
    return alpha * S.ptri * expf(-inc * inc * c.inv2i);
}

template <typename T>
__device__ __forceinline__ T pick8(int k, T v0, T v1, T v2, T v3, T v4, T v5, T v6, T v7)
{
    T r = v0;
    r = k == 1 ? v1 : r;
    r = k == 2 ? v2 : r;
    r = k == 3 ? v3 : r;
    r = k == 4 ? v4 : r;
    r = k == 5 ? v5 : r;
    r = k == 6 ? v6 : r;
    r = k == 7 ? v7 : r;
    return r;
}

// Costs of the plane (d, a, b) against the G sources s0 .. s0+G-1 in one
// shared window pass, stored in s_c0[s][tid].
template <int G>
__device__ __forceinline__ void select_group(const PMConst &c, const SrcLds *s_src, int s0, float du,
                                             float dv, float d, float a, float b, const float *s_ref,
                                             int lbase, const RefStats &R, float *s_c0, int tid,
                                             bool hyp_ok)
{
    Proj P[G];
    float cg[G];
    const SrcLds *SS[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        P[k] = make_proj(s_src[s0 + k], du, dv, d, a, b);
        SS[k] = &s_src[s0 + k];
    }
    ncc_cost_group<G>(c, SS, P, s_ref, lbase, R, cg, nullptr, hyp_ok);
#pragma unroll
    for (int k = 0; k < G; ++k) s_c0[(s0 + k) * kThreads + tid] = cg[k];
}

// Candidate hypotheses K0+1 .. K0+G against source Ss in one shared window
// pass; acc[k] += cnt * cost (photometric, plus the geometric term in pass 1).
template <int G, int K0>
__device__ __forceinline__ void eval_hyp_group(const PMConst &c, const SrcLds &Ss, float cnt, int pass,
                                               float du, float dv, const float (&hd)[kNHyp],
                                               const float (&ak)[kNHyp - 1],
                                               const float (&bk)[kNHyp - 1], const float *s_ref,
                                               int lbase, const RefStats &R, float (&acc)[kNHyp - 1],
                                               bool hyp_ok)
{
    Proj P[G];
    float ph[G];
    const SrcLds *SS[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        P[k] = make_proj(Ss, du, dv, hd[K0 + k + 1], ak[K0 + k], bk[K0 + k]);
        SS[k] = &Ss;
    }
    // geometric term first: its depth gathers overlap the window pass
    float gt[G];
#pragma unroll
    for (int k = 0; k < G; ++k)
        gt[k] = pass == 1 ? c.geom_reg * fminf(geo_error(c, Ss, P[k], du, dv), c.geom_max) : 0.f;
    ncc_cost_group<G>(c, SS, P, s_ref, lbase, R, ph, nullptr, hyp_ok);
#pragma unroll
    for (int k = 0; k < G; ++k) acc[K0 + k] += cnt * (pass == 1 ? ph[k] + gt[k] : ph[k]);
}

// All candidate hypotheses 1 .. kNHyp-1 in groups of (at most) G.
template <int G, int K0>
__device__ __forceinline__ void eval_hyps(const PMConst &c, const SrcLds &Ss, float cnt, int pass,
                                          float du, float dv, const float (&hd)[kNHyp],
                                          const float (&ak)[kNHyp - 1], const float (&bk)[kNHyp - 1],
                                          const float *s_ref, int lbase, const RefStats &R,
                                          float (&acc)[kNHyp - 1], bool hyp_ok)
{
    if constexpr (K0 < kNHyp - 1) {
        constexpr int g = (kNHyp - 1 - K0) < G ? (kNHyp - 1 - K0) : G;
        eval_hyp_group<g, K0>(c, Ss, cnt, pass, du, dv, hd, ak, bk, s_ref, lbase, R, acc, hyp_ok);
        eval_hyps<G, K0 + g>(c, Ss, cnt, pass, du, dv, hd, ak, bk, s_ref, lbase, R, acc, hyp_ok);
    }
}

// Tile shapes: TW x TH pixels with TW * TH = 2 kThreads, so in a red-black
// half-step each thread owns one pixel of the active colour; a wave covers
// 128 / TW rows of the tile.  The reference tile with its halo sits in LDS
// with row pitch TW + 2 kMaxRadius.
template <int TW> struct TileShape {
    static constexpr int W = TW, H = 2 * kThreads / TW;
    static constexpr int LW = TW + 2 * kMaxRadius, LH = H + 2 * kMaxRadius;
};

// Common tile prologue: reference tile (replicated border) into LDS and the
// tile-centred source matrices.
template <int MAXS, int TW>
__device__ __forceinline__ void tile_prologue(const ViewDesc &V, const PMConst &c, const Tile &t,
                                              float *s_ref, SrcLds *s_src)
{
    using T = TileShape<TW>;
    const int W = V.W, H = V.H;
    for (int k = threadIdx.x; k < T::LH * T::LW; k += kThreads) {
        const int ly = k / T::LW, lx = k - ly * T::LW;
        const int gx = min(max(t.x0 + lx - kMaxRadius, 0), W - 1);
        const int gy = min(max(t.y0 + ly - kMaxRadius, 0), H - 1);
        s_ref[k] = (float)V.ref[gy * W + gx];
    }
    if ((int)threadIdx.x < V.n_src)
        setup_source(V, c, threadIdx.x, t.x0 + T::W / 2, t.y0 + T::H / 2, T::W / 2, T::H / 2,
                     s_src[threadIdx.x]);
}

// One red-black half-step on one tile (all 256 threads of the workgroup
// call it; it contains one barrier).
template <int MAXS, int TW>
__device__ __forceinline__ void halfstep_tile(const ViewDesc &V, const Tile &t, const PMConst &c,
                                              float *s_ref, SrcLds *s_src, float *s_c0,
                                              uint32_t *s_list, float *s_w, unsigned int &my_ncc,
                                              unsigned int &my_geo)
{
    using T = TileShape<TW>;
    const int iter = V.step >> 1, color = V.step & 1, pass = V.pass;
    const int W = V.W, H = V.H, S = V.n_src;
    const int tid = threadIdx.x;
    tile_prologue<MAXS, TW>(V, c, t, s_ref, s_src);
    __syncthreads();

    constexpr int kHalf = TW / 2;               // active pixels per tile row
    const int ty = tid / kHalf;
    const int tx = ((tid % kHalf) << 1) + ((t.y0 + ty + color) & 1);
    const int x = t.x0 + tx, y = t.y0 + ty;
    if (x < W && y < H) {
        const size_t HW = (size_t)W * H;
        const int p = y * W + x;
        const float du = (float)(tx - T::W / 2), dv = (float)(ty - T::H / 2);
        const int lbase = (ty + kMaxRadius) * T::LW + tx + kMaxRadius;

        // ---- candidate hypotheses ---------------------------------------
        float hd[kNHyp], hn0[kNHyp], hn1[kNHyp], hn2[kNHyp];
        const float4 h0 = V.hyp[p];
        hd[0] = h0.x; hn0[0] = h0.y; hn1[0] = h0.z; hn2[0] = h0.w;
        const int DX[4] = {-1, 1, 0, 0}, DY[4] = {0, 0, -1, 1};
        // lowest-cost neighbour at odd distances 1, 3, .., 2 n_far - 1 along
        // each direction (the first of equal costs); the candidate costs of a
        // chunk are loaded together, then compared in order
#pragma unroll
        for (int dir = 0; dir < 4; ++dir) {
            int bx = -1, by = -1;
            float bc = 0.f;
            for (int k0 = 0; k0 < c.n_far; k0 += kFarChunk) {
                float cq[kFarChunk];
#pragma unroll
                for (int kk = 0; kk < kFarChunk; ++kk) {
                    const int k = k0 + kk, dist = 2 * k + 1;
                    const int qx = x + DX[dir] * dist, qy = y + DY[dir] * dist;
                    const bool ok = k < c.n_far && qx >= 0 && qy >= 0 && qx < W && qy < H;
                    cq[kk] = ok ? V.cost[qy * W + qx] : 0.f;
                }
#pragma unroll
                for (int kk = 0; kk < kFarChunk; ++kk) {
                    const int k = k0 + kk, dist = 2 * k + 1;
                    const int qx = x + DX[dir] * dist, qy = y + DY[dir] * dist;
                    const bool ok = k < c.n_far && qx >= 0 && qy >= 0 && qx < W && qy < H;
                    if (ok && (bx < 0 || cq[kk] < bc)) { bx = qx; by = qy; bc = cq[kk]; }
                }
            }
            if (bx < 0) {
                hd[1 + dir] = hd[0]; hn0[1 + dir] = hn0[0]; hn1[1 + dir] = hn1[0]; hn2[1 + dir] = hn2[0];
            } else {
                const float4 hq = V.hyp[by * W + bx];
                hd[1 + dir] = hq.x - (hq.y * (float)(x - bx) + hq.z * (float)(y - by)) / hq.w;
                hn0[1 + dir] = hq.y; hn1[1 + dir] = hq.z; hn2[1 + dir] = hq.w;
            }
        }
        const uint32_t code = step_code(pass, iter, color);
        float u0[4], u1[4];
        rng_block(c, V.view_id, (uint32_t)p, code, 0u, u0);
        rng_block(c, V.view_id, (uint32_t)p, code, 1u, u1);
        const float range = V.depth_max - V.depth_min;
        const float decay = powf(c.decay_base, (float)iter);
        {
            float n[3];
            hd[5] = V.depth_min + u0[0] * range;
            random_normal(c, u0[1], u0[2], n);
            hn0[5] = n[0]; hn1[5] = n[1]; hn2[5] = n[2];
            float d6 = hd[0] + (2.f * u0[3] - 1.f) * c.perturbation * range * decay;
            hd[6] = fminf(fmaxf(d6, V.depth_min), V.depth_max);
            hn0[6] = hn0[0]; hn1[6] = hn1[0]; hn2[6] = hn2[0];
            const float nn[3] = {hn0[0], hn1[0], hn2[0]};
            hd[7] = hd[0];
            perturb_normal(c, nn, c.normal_perturbation * decay, u1[0], u1[1], n);
            hn0[7] = n[0]; hn1[7] = n[1]; hn2[7] = n[2];
        }

        // every candidate within the depth and slope bounds of the tile's
        // `inner` sources (|a| = |n0 / n2| <= amax checked without dividing,
        // with a margin for the division's rounding)
        bool hyp_ok = true;
        {
            const float amx = 0.999f * slope_bound(c), dlo = hyp_depth_lo(V), dhi = hyp_depth_hi(V);
#pragma unroll
            for (int k = 0; k < kNHyp; ++k)
                hyp_ok = hyp_ok && hd[k] >= dlo && hd[k] <= dhi && hn2[k] > 0.f &&
                         fabsf(hn0[k]) <= amx * hn2[k] && fabsf(hn1[k]) <= amx * hn2[k];
        }

        // ---- view selection on the current hypothesis --------------------
        const bool wtab = c.ntaps * c.ntaps <= kWTab;
        const RefStats R = ref_stats(c, s_ref, lbase, T::LW, wtab ? s_w + tid : nullptr);
        const float a0 = -hn0[0] / hn2[0], b0 = -hn1[0] / hn2[0];
        const float nh0[3] = {hn0[0], hn1[0], hn2[0]};
        float jtn;
        {
            const float *J = V.J;
            const float j0 = J[0] * nh0[0] + J[3] * nh0[1] + J[6] * nh0[2];
            const float j1 = J[1] * nh0[0] + J[4] * nh0[1] + J[7] * nh0[2];
            const float j2 = J[2] * nh0[0] + J[5] * nh0[1] + J[8] * nh0[2];
            jtn = sqrtf(j0 * j0 + j1 * j1 + j2 * j2);
        }
        // costs of the current plane against every source, up to kGroupV
        // sources sharing one window pass
        for (int s0 = 0; s0 < S;) {
            const int left = S - s0;
            if (kGroupV >= 4 && (left >= 7 || left == 4 || left == 5)) {
                select_group<4>(c, s_src, s0, du, dv, hd[0], a0, b0, s_ref, lbase, R, s_c0, tid, hyp_ok);
                s0 += 4;
            } else if (kGroupV >= 3 && (left == 3 || left == 6)) {
                select_group<3>(c, s_src, s0, du, dv, hd[0], a0, b0, s_ref, lbase, R, s_c0, tid, hyp_ok);
                s0 += 3;
            } else if (kGroupV >= 2 && left >= 2) {
                select_group<2>(c, s_src, s0, du, dv, hd[0], a0, b0, s_ref, lbase, R, s_c0, tid, hyp_ok);
                s0 += 2;
            } else {
                select_group<1>(c, s_src, s0, du, dv, hd[0], a0, b0, s_ref, lbase, R, s_c0, tid, hyp_ok);
                s0 += 1;
            }
        }
        float tot = 0.f;
        int lastpos = 0;
        bool inb[4];
#pragma unroll
        for (int dir = 0; dir < 4; ++dir) {
            const int qx = x + DX[dir], qy = y + DY[dir];
            inb[dir] = qx >= 0 && qy >= 0 && qx < W && qy < H;
        }
        for (int s0 = 0; s0 < S; s0 += kProbChunk) {
          // the neighbours' selection probabilities of a chunk of sources,
          // loaded together
          float pn[kProbChunk][4];
#pragma unroll
          for (int kk = 0; kk < kProbChunk; ++kk)
#pragma unroll
              for (int dir = 0; dir < 4; ++dir)
                  pn[kk][dir] = (s0 + kk < S && inb[dir])
                                    ? V.prob[(size_t)(s0 + kk) * HW + (y + DY[dir]) * W + (x + DX[dir])]
                                    : 0.f;
          float c0v[kProbChunk];
#pragma unroll
          for (int kk = 0; kk < kProbChunk; ++kk)
              c0v[kk] = s0 + kk < S ? s_c0[(s0 + kk) * kThreads + tid] : 0.f;
#pragma unroll
          for (int kk = 0; kk < kProbChunk; ++kk) {
            const int s = s0 + kk;
            if (s >= S) break;
            const SrcLds &Ss = s_src[s];
            const float c0 = c0v[kk];
            float prev = 0.f;
            int np = 0;
#pragma unroll
            for (int dir = 0; dir < 4; ++dir) {
                if (!inb[dir]) continue;
                prev += pn[kk][dir];
                ++np;
            }
            prev = np ? prev / (float)np : 0.5f;
            const float L = expf(-c0 * c0 * c.inv2n);
            const float a1 = L * (prev * 0.999f + (1.f - prev) * 0.001f);
            const float a0v = 0.5f * ((1.f - prev) * 0.999f + prev * 0.001f);
            const float alpha = a1 / (a1 + a0v);
            V.prob[(size_t)s * HW + p] = alpha;
            const float q = view_weight(c, Ss, alpha, nh0, jtn);
            // q waits in the sampled-list slot of source s: the draw below
            // reads slot s before it writes list entry nlist <= s
            s_list[s * kThreads + tid] = __float_as_uint(q);
            tot += q;
            if (q > 0.f) lastpos = s;
          }
        }
        my_ncc += S;

        // ---- Monte-Carlo draw of sources ---------------------------------
        int nlist = 0;
        float wsum = 0.f, acc0 = 0.f;
        if (!(tot > 0.f)) {
            for (int s = 0; s < S; ++s) {
                const SrcLds &Ss = s_src[s];
                float cc = s_c0[s * kThreads + tid];
                if (pass == 1) {
                    const Proj P = make_proj(Ss, du, dv, hd[0], a0, b0);
                    cc = total_cost(c, Ss, P, cc, pass, du, dv);
                    ++my_geo;
                }
                acc0 += 1.f * cc;
                wsum += 1.f;
                s_list[s * kThreads + tid] = 1u;
                ++nlist;
            }
        } else {
            float tk[kMaxSamples];
            // the draws' counter word: the pixel, or its group of
            // g x 128 / g pixels (g = c.draw_group, the tile width: the
            // pixels one wave covers), which share the uniforms
            uint32_t dkey = (uint32_t)p;
            if (c.draw_group) {
                const int gw = c.draw_group, gh = 128 / gw;
                dkey = (uint32_t)((y / gh) * ((W + gw - 1) / gw) + x / gw);
            }
#pragma unroll
            for (int b = 0; b < kMaxSamples / 4; ++b) {
                float u[4];
                rng_block(c, V.view_id, dkey, code, 2u + (uint32_t)b, u);
                tk[4 * b + 0] = u[0] * tot; tk[4 * b + 1] = u[1] * tot;
                tk[4 * b + 2] = u[2] * tot; tk[4 * b + 3] = u[3] * tot;
            }
            // samples beyond num_samples are never counted
#pragma unroll
            for (int k = 0; k < kMaxSamples; ++k)
                if (k >= c.num_samples) tk[k] = INFINITY;
            // Source s takes the samples in [lo_s, hi_s): lo_s is the sum of
            // the weights before s, hi_s = lo_s + q_s, and the last source of
            // positive weight takes everything from its lo up (hi = inf for
            // it and for every later source, which then counts 0).  The count
            // is #(t < hi_s) - #(t < lo_s); #(t < lo_s) is the previous
            // source's #(t < hi).
            float acc = 0.f;
            int below = 0;
            for (int s = 0; s < S; ++s) {
                const SrcLds &Ss = s_src[s];
                const float q = __uint_as_float(s_list[s * kThreads + tid]);
                const float c0s = s_c0[s * kThreads + tid];
                acc = acc + q;
                const float hi = s >= lastpos ? INFINITY : acc;
                int nb = 0;
#pragma unroll
                for (int k = 0; k < kMaxSamples; ++k) nb += tk[k] < hi ? 1 : 0;
                const int cnt = nb - below;
                below = nb;
                if (cnt > 0) {
                    float cc = c0s;
                    if (pass == 1) {
                        const Proj P = make_proj(Ss, du, dv, hd[0], a0, b0);
                        cc = total_cost(c, Ss, P, cc, pass, du, dv);
                        ++my_geo;
                    }
                    acc0 += (float)cnt * cc;
                    wsum += (float)cnt;
                    ++nlist;
                }
                // slot s now holds the draw count of source s (0: not drawn)
                s_list[s * kThreads + tid] = (uint32_t)cnt;
            }
        }

        // ---- evaluate the other seven hypotheses -------------------------
        // One pass over the window per sampled source evaluates all seven.
        // Sources are visited in ascending order by the whole wave (a lane
        // that did not draw source s is masked off), so every gather of a
        // pass reads one source image; each lane still accumulates its drawn
        // sources in ascending order.
        float ak[kNHyp - 1], bk[kNHyp - 1], acc[kNHyp - 1];
#pragma unroll
        for (int k = 0; k < kNHyp - 1; ++k) {
            ak[k] = -hn0[k + 1] / hn2[k + 1];
            bk[k] = -hn1[k + 1] / hn2[k + 1];
            acc[k] = 0.f;
        }
        for (int s = 0; s < S; ++s) {
            const uint32_t e = s_list[s * kThreads + tid];
            if (e == 0u) continue;
            const SrcLds &Ss = s_src[s];
            const float cnt = (float)e;
            eval_hyps<kGroupH, 0>(c, Ss, cnt, pass, du, dv, hd, ak, bk, s_ref, lbase, R, acc, hyp_ok);
        }
        my_ncc += (kNHyp - 1) * nlist;
        if (pass == 1) my_geo += (kNHyp - 1) * nlist;
        float best_c = acc0 / wsum;
        int best = 0;
#pragma unroll
        for (int k = 0; k < kNHyp - 1; ++k) {
            const float ck = acc[k] / wsum;
            if (ck < best_c) { best_c = ck; best = k + 1; }
        }
        const float bd = pick8(best, hd[0], hd[1], hd[2], hd[3], hd[4], hd[5], hd[6], hd[7]);
        const float b0n = pick8(best, hn0[0], hn0[1], hn0[2], hn0[3], hn0[4], hn0[5], hn0[6], hn0[7]);
        const float b1n = pick8(best, hn1[0], hn1[1], hn1[2], hn1[3], hn1[4], hn1[5], hn1[6], hn1[7]);
        const float b2n = pick8(best, hn2[0], hn2[1], hn2[2], hn2[3], hn2[4], hn2[5], hn2[6], hn2[7]);
        V.hyp[p] = make_float4(bd, b0n, b1n, b2n);
        V.cost[p] = best_c;
    }
}

constexpr int kXcd = 8;   // MI355X: 8 XCDs, workgroups dispatched round-robin across them

// Persistent red-black half-step over a tile list.  Workgroup b runs on XCD
// b % 8 and takes tiles from that XCD's contiguous share of the list (the
// views' source images stay in that XCD's L2), then steals from the other
// XCDs' shares.  work[8] are per-XCD cursors zeroed before the launch.  The
// order tiles are taken in never changes a result: tiles are independent
// within a half-step.
template <int MAXS, int TW>
__global__ __launch_bounds__(kThreads, 2) void k_halfstep(const ViewDesc *__restrict__ views,
                                                       const Tile *__restrict__ tiles, int ntiles,
                                                       PMConst c, Counters *counters,
                                                       unsigned int *work, float *c0g)
{
    __shared__ float s_ref[TileShape<TW>::LH * TileShape<TW>::LW];
    __shared__ SrcLds s_src[MAXS];
    // kernels for more than 16 sources keep the view-selection costs in a
    // per-workgroup global scratch, which leaves LDS room for the tap weights
    __shared__ float s_c0[MAXS > 16 ? 1 : MAXS * kThreads];
    __shared__ uint32_t s_list[MAXS * kThreads];
    __shared__ float s_w[kWTab * kThreads];
    __shared__ unsigned int s_evals[2];
    __shared__ int s_tile;
    const int tid = threadIdx.x;
    if (tid < 2) s_evals[tid] = 0u;
    unsigned int my_ncc = 0, my_geo = 0;
    const int chunk = (ntiles + kXcd - 1) / kXcd;
    const int home = blockIdx.x % kXcd;
    for (int probe = 0; probe < kXcd; ++probe) {
        const int xcd = (home + probe) % kXcd;
        const int lo = xcd * chunk, hi = min(ntiles, lo + chunk);
        while (true) {
            if (tid == 0) {
                const int i = (int)atomicAdd(&work[xcd], 1u);
                s_tile = lo + i < hi ? lo + i : -1;
            }
            __syncthreads();
            const int ti = s_tile;
            __syncthreads();
            if (ti < 0) break;
            const Tile t = tiles[ti];
            const ViewDesc &V = views[t.view];
            if (!V.active || V.step >= V.num_steps) continue;     // uniform per workgroup
            halfstep_tile<MAXS, TW>(V, t, c, s_ref, s_src,
                                MAXS > 16 ? c0g + (size_t)blockIdx.x * (MAXS * kThreads) : s_c0,
                                s_list, s_w, my_ncc, my_geo);
        }
    }
    if (counters) {
        atomicAdd(&s_evals[0], my_ncc);
        atomicAdd(&s_evals[1], my_geo);
        __syncthreads();
        if (tid == 0) {
            atomicAdd(&counters->ncc_evals, (unsigned long long)s_evals[0]);
            atomicAdd(&counters->geo_evals, (unsigned long long)s_evals[1]);
        }
    }
}

// Random initialisation: uniform depth in the range, normal in the cone
// (tiles of tw x 2 kThreads / tw pixels).
__global__ __launch_bounds__(kThreads) void k_init(const ViewDesc *__restrict__ views,
                                                   const Tile *__restrict__ tiles, PMConst c,
                                                   int random_hyp, int tw)
{
    const Tile t = tiles[blockIdx.x];
    const ViewDesc &V = views[t.view];
    const size_t HW = (size_t)V.W * V.H;
    for (int k = threadIdx.x; k < 2 * kThreads; k += kThreads) {
        const int x = t.x0 + k % tw, y = t.y0 + k / tw;
        if (x >= V.W || y >= V.H) continue;
        const int p = y * V.W + x;
        if (random_hyp) {
            float u[4], n[3];
            rng_block(c, V.view_id, (uint32_t)p, init_code(V.pass), 0u, u);
            random_normal(c, u[1], u[2], n);
            V.hyp[p] = make_float4(V.depth_min + u[0] * (V.depth_max - V.depth_min), n[0], n[1], n[2]);
        }
        V.cost[p] = kMaxCost;
        for (int s = 0; s < V.n_src; ++s) V.prob[(size_t)s * HW + p] = 0.5f;
    }
}

// Cost of the stored hypotheses against every source: out[s*HW + p].
template <int MAXS>
__global__ __launch_bounds__(kThreads) void k_eval_cost(const ViewDesc *__restrict__ views,
                                                        const Tile *__restrict__ tiles, PMConst c,
                                                        float *out)
{
    __shared__ float s_ref[kLdsH * kLdsW];
    __shared__ SrcLds s_src[MAXS];
    const Tile t = tiles[blockIdx.x];
    const ViewDesc &V = views[t.view];
    tile_prologue<MAXS, kTileW>(V, c, t, s_ref, s_src);
    __syncthreads();
    const size_t HW = (size_t)V.W * V.H;
    for (int k = threadIdx.x; k < kTileW * kTileH; k += kThreads) {
        const int tx = k & (kTileW - 1), ty = k / kTileW;
        const int x = t.x0 + tx, y = t.y0 + ty;
        if (x >= V.W || y >= V.H) continue;
        const int p = y * V.W + x;
        const int lbase = (ty + kMaxRadius) * kLdsW + tx + kMaxRadius;
        const float du = (float)(tx - kTileW / 2), dv = (float)(ty - kTileH / 2);
        const RefStats R = ref_stats(c, s_ref, lbase, kLdsW);
        const float4 h = V.hyp[p];
        const float a = -h.y / h.w, b = -h.z / h.w;
        for (int s = 0; s < V.n_src; ++s) {
            const SrcLds &Ss = s_src[s];
            const Proj P = make_proj(Ss, du, dv, h.x, a, b);
            const float ph = ncc_cost(c, Ss, P, s_ref, lbase, R);
            out[(size_t)s * HW + p] = total_cost(c, Ss, P, ph, V.pass, du, dv);
        }
    }
}

// ---------------------------------------------------------------------------
// LDS staging of source footprints (measurement kernel).  Cost of the stored
// hypotheses against every source, sources outer: per (tile, source) the
// workgroup reduces the bounding box of every window's projection (the
// projective image of a square window is the convex hull of its corners'
// images; +1 texel for the bilinear neighbour, clamped to the image), and
// when the box has at most kStageCap texels it copies the box's quads (words
// of the pair texture) into LDS with row-contiguous loads, after which every
// tap reads LDS instead of the L1/L2-cached global texture.  STAGE = false runs the
// same loops with global gathers only, so the two variants time the gather
// path alone; their outputs are identical (same texels, same arithmetic).
// ---------------------------------------------------------------------------
constexpr int kStageCap = 4096;      // texels: 16 KB of LDS

__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

template <int MAXS, bool STAGE>
__global__ __launch_bounds__(kThreads) void k_eval_stage(const ViewDesc *__restrict__ views,
                                                         const Tile *__restrict__ tiles, PMConst c,
                                                         float *out, unsigned int *n_staged)
{
    __shared__ float s_ref[kLdsH * kLdsW];
    __shared__ SrcLds s_src[MAXS];
    __shared__ uint32_t s_stage[STAGE ? kStageCap : 1];
    __shared__ int s_box[4];
    const Tile t = tiles[blockIdx.x];
    const ViewDesc &V = views[t.view];
    tile_prologue<MAXS, kTileW>(V, c, t, s_ref, s_src);
    __syncthreads();
    const size_t HW = (size_t)V.W * V.H;
    constexpr int kPix = kTileW * kTileH / kThreads;          // pixels per thread
    bool inb[kPix];
    int pix[kPix], lb[kPix];
    float du[kPix], dv[kPix], hd[kPix], ha[kPix], hb[kPix];
    RefStats R[kPix];
#pragma unroll
    for (int j = 0; j < kPix; ++j) {
        const int k = threadIdx.x + j * kThreads;
        const int tx = k & (kTileW - 1), ty = k / kTileW;
        const int x = t.x0 + tx, y = t.y0 + ty;
        inb[j] = x < V.W && y < V.H;
        pix[j] = inb[j] ? y * V.W + x : 0;
        lb[j] = (ty + kMaxRadius) * kLdsW + tx + kMaxRadius;
        du[j] = (float)(tx - kTileW / 2);
        dv[j] = (float)(ty - kTileH / 2);
        R[j] = ref_stats(c, s_ref, lb[j], kLdsW);
        const float4 h = V.hyp[pix[j]];
        hd[j] = h.x;
        ha[j] = -h.y / h.w;
        hb[j] = -h.z / h.w;
    }
    const float fr = (float)c.radius;
    for (int s = 0; s < V.n_src; ++s) {
        const SrcLds &Ss = s_src[s];
        bool staged = false;
        SrcView lv;
        if constexpr (STAGE) {
            // this thread's windows: bounding box in anchor-relative texels
            int x0 = INT_MAX, y0 = INT_MAX, x1 = INT_MIN, y1 = INT_MIN;
            bool unbounded = false;
#pragma unroll
            for (int j = 0; j < kPix; ++j) {
                if (!inb[j]) continue;
                const Proj P = make_proj(Ss, du[j], dv[j], hd[j], ha[j], hb[j]);
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) {
                    const float ox = (cc & 1) ? fr : -fr, oy = (cc & 2) ? fr : -fr;
                    const float qx = fmaf(oy, P.hvx, fmaf(ox, P.hux, P.q0x));
                    const float qy = fmaf(oy, P.hvy, fmaf(ox, P.huy, P.q0y));
                    const float qz = fmaf(oy, P.hvz, fmaf(ox, P.huz, P.q0z));
                    if (!(qz > 0.f)) { unbounded = true; continue; }
                    const float px = __builtin_amdgcn_fmed3f(qx / qz, Ss.lox, Ss.hix);
                    const float py = __builtin_amdgcn_fmed3f(qy / qz, Ss.loy, Ss.hiy);
                    x0 = min(x0, (int)floorf(px) - 1);
                    y0 = min(y0, (int)floorf(py) - 1);
                    x1 = max(x1, (int)floorf(px) + 2);
                    y1 = max(y1, (int)floorf(py) + 2);
                }
            }
            if (unbounded) { x0 = INT_MIN / 2; y0 = INT_MIN / 2; x1 = INT_MAX / 2; y1 = INT_MAX / 2; }
            if (threadIdx.x == 0) { s_box[0] = INT_MAX; s_box[1] = INT_MAX; s_box[2] = INT_MIN; s_box[3] = INT_MIN; }
            __syncthreads();
            x0 = wave_min(x0); y0 = wave_min(y0); x1 = wave_max(x1); y1 = wave_max(y1);
            if ((threadIdx.x & 63) == 0) {
                atomicMin(&s_box[0], x0); atomicMin(&s_box[1], y0);
                atomicMax(&s_box[2], x1); atomicMax(&s_box[3], y1);
            }
            __syncthreads();
            // absolute texel box, clamped to the image (clamped taps stay inside it)
            const int bx0 = max(s_box[0] + Ss.ax, 0), by0 = max(s_box[1] + Ss.ay, 0);
            const int bx1 = min(s_box[2] + Ss.ax, Ss.W - 1), by1 = min(s_box[3] + Ss.ay, Ss.H - 1);
            const int bw = bx1 - bx0 + 1, bh = by1 - by0 + 1;
            staged = bw > 0 && bh > 0 && (long long)bw * bh <= kStageCap;
            if (staged) {
                // row-contiguous copy of the box
                int xx = threadIdx.x % bw, yy = threadIdx.x / bw;
                for (int i = threadIdx.x; i < bw * bh; i += kThreads) {
                    s_stage[i] = *(gword_t)((gbyte_t)Ss.pair + 2 * ((size_t)(by0 + yy) * Ss.pitch + bx0 + xx));
                    xx += kThreads % bw;
                    yy += kThreads / bw;
                    if (xx >= bw) { xx -= bw; ++yy; }
                }
                lv = view_of(Ss);
                lv.pitchb = bw;
                lv.baseb = (Ss.ay - by0) * bw + (Ss.ax - bx0);
                lv.pair = (const uint8_t *)s_stage;
                if (threadIdx.x == 0 && n_staged) atomicAdd(n_staged, 1u);
            }
            __syncthreads();
        }
#pragma unroll
        for (int j = 0; j < kPix; ++j) {
            const Proj P = make_proj(Ss, du[j], dv[j], hd[j], ha[j], hb[j]);
            const Proj PP[1] = {P};
            const SrcLds *const SS[1] = {&Ss};
            float cost[1];
            if (STAGE && staged)
                ncc_cost_group<1, true>(c, SS, PP, s_ref, lb[j], R[j], cost, &lv);
            else
                ncc_cost_group<1>(c, SS, PP, s_ref, lb[j], R[j], cost);
            if (inb[j]) out[(size_t)s * HW + pix[j]] = total_cost(c, Ss, P, cost[0], V.pass, du[j], dv[j]);
        }
        if constexpr (STAGE) __syncthreads();     // the next source's copy overwrites the box
    }
}

// Filter rule of the plane (d, a, b) against the G sources s0 .. s0+G-1,
// their costs from one shared window pass: a source is consistent when
// 1 - cost >= min NCC, its triangulation angle passes and, in the geometric
// pass, the forward-backward error passes; bit s of m is set for it.
template <int G>
__device__ __forceinline__ void filter_group(const PMConst &c, int pass, const SrcLds *s_src, int s0,
                                             float du, float dv, float d, float a, float b,
                                             const float *s_ref, int lbase, const RefStats &R,
                                             uint32_t &m, int &count)
{
    Proj P[G];
    float ph[G];
    const SrcLds *SS[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        P[k] = make_proj(s_src[s0 + k], du, dv, d, a, b);
        SS[k] = &s_src[s0 + k];
    }
    ncc_cost_group<G>(c, SS, P, s_ref, lbase, R, ph);
#pragma unroll
    for (int k = 0; k < G; ++k) {
        const SrcLds &Ss = s_src[s0 + k];
        if (!(1.f - ph[k] >= c.f_min_ncc)) continue;
        if (!(Ss.tri >= c.f_min_tri)) continue;
        if (pass == 1 && !(geo_error(c, Ss, P[k], du, dv) <= c.f_geom_max)) continue;
        m |= 1u << (s0 + k);
        ++count;
    }
}

// Filter: count consistent sources; zero the hypothesis when too few.
template <int MAXS>
__global__ __launch_bounds__(kThreads) void k_filter(const ViewDesc *__restrict__ views,
                                                     const Tile *__restrict__ tiles, PMConst c,
                                                     uint32_t *mask_out)
{
    __shared__ float s_ref[kLdsH * kLdsW];
    __shared__ SrcLds s_src[MAXS];
    __shared__ float s_w[kWTab * kThreads];
    const Tile t = tiles[blockIdx.x];
    const ViewDesc &V = views[t.view];
    const bool wtab = c.ntaps * c.ntaps <= kWTab;
    tile_prologue<MAXS, kTileW>(V, c, t, s_ref, s_src);
    __syncthreads();
    for (int k = threadIdx.x; k < kTileW * kTileH; k += kThreads) {
        const int tx = k & (kTileW - 1), ty = k / kTileW;
        const int x = t.x0 + tx, y = t.y0 + ty;
        if (x >= V.W || y >= V.H) continue;
        const int p = y * V.W + x;
        const float4 h = V.hyp[p];
        uint32_t m = 0u;
        int count = 0;
        if (!(h.x > 0.f)) {
            if (mask_out) mask_out[p] = 0u;
            continue;
        }
        {
            const int lbase = (ty + kMaxRadius) * kLdsW + tx + kMaxRadius;
            const float du = (float)(tx - kTileW / 2), dv = (float)(ty - kTileH / 2);
            const RefStats R = ref_stats(c, s_ref, lbase, kLdsW, wtab ? s_w + threadIdx.x : nullptr);
            const float a = -h.y / h.w, b = -h.z / h.w;
            // up to four sources share a window pass (as in view selection)
            const int S = V.n_src;
            for (int s0 = 0; s0 < S;) {
                const int left = S - s0;
                if (left >= 7 || left == 4 || left == 5) {
                    filter_group<4>(c, V.pass, s_src, s0, du, dv, h.x, a, b, s_ref, lbase, R, m, count);
                    s0 += 4;
                } else if (left == 3 || left == 6) {
                    filter_group<3>(c, V.pass, s_src, s0, du, dv, h.x, a, b, s_ref, lbase, R, m, count);
                    s0 += 3;
                } else if (left == 2) {
                    filter_group<2>(c, V.pass, s_src, s0, du, dv, h.x, a, b, s_ref, lbase, R, m, count);
                    s0 += 2;
                } else {
                    filter_group<1>(c, V.pass, s_src, s0, du, dv, h.x, a, b, s_ref, lbase, R, m, count);
                    s0 += 1;
                }
            }
        }
        if (mask_out) mask_out[p] = m;
        if (count < c.f_min_consistent) V.hyp[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// Outputs: depth, camera-frame normal (facing the camera), cost.
__global__ void k_finalize(const ViewDesc *__restrict__ views, int view, float *depth,
                           float *normal, float *cost_out, float *depth_keep)
{
    const ViewDesc &V = views[view];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= V.W * V.H) return;
    const float4 h = V.hyp[p];
    if (depth) depth[p] = h.x;
    if (depth_keep) depth_keep[p] = h.x;
    if (cost_out) cost_out[p] = V.cost[p];
    if (normal) {
        const float *N = V.N_out;
        float n0 = N[0] * h.y + N[1] * h.z + N[2] * h.w;
        float n1 = N[3] * h.y + N[4] * h.z + N[5] * h.w;
        float n2 = N[6] * h.y + N[7] * h.z + N[8] * h.w;
        const float l = sqrtf(n0 * n0 + n1 * n1 + n2 * n2);
        if (l > 0.f) { n0 /= l; n1 /= l; n2 /= l; }
        if (n2 > 0.f) { n0 = -n0; n1 = -n1; n2 = -n2; }
        normal[3 * p + 0] = n0; normal[3 * p + 1] = n1; normal[3 * p + 2] = n2;
    }
}

// Hypotheses from depth + camera-frame normals (resume from workspace files).
__global__ void k_load_hyp(const ViewDesc *__restrict__ views, int view, const float *depth,
                           const float *normal, float c_cos_max)
{
    const ViewDesc &V = views[view];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= V.W * V.H) return;
    float d = depth[p];
    const float *N = V.N_in;
    const float c0 = normal[3 * p], c1 = normal[3 * p + 1], c2 = normal[3 * p + 2];
    float n0 = N[0] * c0 + N[1] * c1 + N[2] * c2;
    float n1 = N[3] * c0 + N[4] * c1 + N[5] * c2;
    float n2 = N[6] * c0 + N[7] * c1 + N[8] * c2;
    float l = sqrtf(n0 * n0 + n1 * n1 + n2 * n2);
    if (!(d > 0.f) || !(l > 0.f)) {
        d = 0.5f * (V.depth_min + V.depth_max);
        n0 = 0.f; n1 = 0.f; n2 = 1.f; l = 1.f;
    }
    n0 /= l; n1 /= l; n2 /= l;
    if (n2 < 0.f) { n0 = -n0; n1 = -n1; n2 = -n2; }
    // keep loaded normals inside the cone the solver samples from, so plane
    // slopes (and every projected coordinate) stay finite
    if (!(n2 >= c_cos_max)) {
        const float t = sqrtf(fmaxf(0.f, 1.f - c_cos_max * c_cos_max));
        const float h = sqrtf(n0 * n0 + n1 * n1);
        n0 = h > 0.f ? n0 / h * t : t; n1 = h > 0.f ? n1 / h * t : 0.f; n2 = c_cos_max;
    }
    V.hyp[p] = make_float4(d, n0, n1, n2);
}

// Device-side schedule step: every active view advances one half-step.
__global__ void k_advance(ViewDesc *views, int nviews)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nviews && views[v].active && views[v].step < views[v].num_steps) views[v].step += 1;
}

// Pair texture: W + 1 texels of 2 bytes per row; texel (x, y) holds I(x', y)
// and I(x', y1) with x' = min(x, W-1), y1 = min(y+1, H-1).  The word at texel
// x < W is the bilinear quad of (x, y) with edge-clamped neighbours (bytes
// I(x,y), I(x,y1), I(x1,y), I(x1,y1), x1 = min(x+1, W-1)): one 4-byte gather
// per tap from 2 bytes per pixel.
__global__ void k_pack_pair(const uint8_t *__restrict__ gray, int W, int H, uint16_t *__restrict__ pair)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x > W || y >= H) return;
    const int xs = min(x, W - 1), y1 = min(y + 1, H - 1);
    const uint32_t a = gray[(size_t)y * W + xs], c = gray[(size_t)y1 * W + xs];
    pair[(size_t)y * (W + 1) + x] = (uint16_t)(a | (c << 8));
}

// Depth texture for the geometric term: texel (x,y) holds d(x,y), d(x1,y),
// d(x,y1), d(x1,y1) with x1 = min(x+1, W-1), y1 = min(y+1, H-1).
__global__ void k_pack_depth4(const float *__restrict__ d, int W, int H, float4 *__restrict__ out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W || y >= H) return;
    const int x1 = min(x + 1, W - 1), y1 = min(y + 1, H - 1);
    out[(size_t)y * W + x] = make_float4(d[(size_t)y * W + x], d[(size_t)y * W + x1],
                                         d[(size_t)y1 * W + x], d[(size_t)y1 * W + x1]);
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
// The kernel instantiation for (sources, tile width); null for an
// unsupported width.
typedef void (*halfstep_fn)(const ViewDesc *, const Tile *, int, PMConst, Counters *, unsigned int *,
                            float *);

template <int MAXS>
static halfstep_fn halfstep_for_width(int tw)
{
    switch (tw) {
    case 16: return k_halfstep<MAXS, 16>;
    case 32: return k_halfstep<MAXS, 32>;
    default: return tw == 64 ? k_halfstep<MAXS, 64> : nullptr;
    }
}

// 64-wide tiles need 1.5 KB more LDS for the reference tile than the kernel
// for more than 16 sources has at 2 workgroups per CU: that kernel has 16-
// and 32-wide tiles only.
static halfstep_fn halfstep_kernel(int maxs, int tw)
{
    if (maxs <= 8) return halfstep_for_width<8>(tw);
    if (maxs <= 16) return halfstep_for_width<16>(tw);
    if (tw == 16) return k_halfstep<32, 16>;
    if (tw == 32) return k_halfstep<32, 32>;
    return nullptr;
}

bool tile_width_supported(int tw, int maxs) { return tw == 16 || tw == 32 || (tw == 64 && maxs <= 16); }

hipError_t launch_halfstep(int maxs, int tw, const ViewDesc *views, const Tile *tiles, int ntiles,
                           const PMConst &c, Counters *counters, unsigned int *work, int grid,
                           float *c0_scratch, hipStream_t st)
{
    if (ntiles <= 0) return hipSuccess;
    const halfstep_fn fn = halfstep_kernel(maxs, tw);
    if (!fn) return hipErrorInvalidValue;
    // more than 16 sources: c0_scratch holds grid x 32 x kThreads floats
    if (maxs > 16 && !c0_scratch) return hipErrorInvalidValue;
    hipError_t e = launch_fill_u32(work, kXcd, 0u, st);
    if (e != hipSuccess) return e;
    grid = max(kXcd, min(grid, ((ntiles + kXcd - 1) / kXcd) * kXcd));
    fn<<<grid, kThreads, 0, st>>>(views, tiles, ntiles, c, counters, work, c0_scratch);
    return hipGetLastError();
}

int halfstep_grid(int maxs, int tw)
{
    int dev = 0, ncu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) ncu = 256;
    const halfstep_fn fn = halfstep_kernel(maxs, tw);
    hipError_t e = fn ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kThreads, 0)
                      : hipErrorInvalidValue;
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    return ((ncu * per_cu + kXcd - 1) / kXcd) * kXcd;
}

hipError_t launch_init(const ViewDesc *views, const Tile *tiles, int ntiles, const PMConst &c,
                       int random_hyp, int tw, hipStream_t st)
{
    if (ntiles <= 0) return hipSuccess;
    k_init<<<ntiles, kThreads, 0, st>>>(views, tiles, c, random_hyp, tw);
    return hipGetLastError();
}

hipError_t launch_eval_cost(int maxs, const ViewDesc *views, const Tile *tiles, int ntiles,
                            const PMConst &c, float *out, hipStream_t st)
{
    if (ntiles <= 0) return hipSuccess;
    if (maxs <= 8) k_eval_cost<8><<<ntiles, kThreads, 0, st>>>(views, tiles, c, out);
    else if (maxs <= 16) k_eval_cost<16><<<ntiles, kThreads, 0, st>>>(views, tiles, c, out);
    else k_eval_cost<32><<<ntiles, kThreads, 0, st>>>(views, tiles, c, out);
    return hipGetLastError();
}

hipError_t launch_eval_stage(int maxs, int stage, const ViewDesc *views, const Tile *tiles, int ntiles,
                             const PMConst &c, float *out, unsigned int *n_staged, hipStream_t st)
{
    if (ntiles <= 0) return hipSuccess;
    if (stage) {
        if (maxs <= 8) k_eval_stage<8, true><<<ntiles, kThreads, 0, st>>>(views, tiles, c, out, n_staged);
        else if (maxs <= 16) k_eval_stage<16, true><<<ntiles, kThreads, 0, st>>>(views, tiles, c, out, n_staged);
        else k_eval_stage<32, true><<<ntiles, kThreads, 0, st>>>(views, tiles, c, out, n_staged);
    } else {
        if (maxs <= 8) k_eval_stage<8, false><<<ntiles, kThreads, 0, st>>>(views, tiles, c, out, n_staged);
        else if (maxs <= 16) k_eval_stage<16, false><<<ntiles, kThreads, 0, st>>>(views, tiles, c, out, n_staged);
        else k_eval_stage<32, false><<<ntiles, kThreads, 0, st>>>(views, tiles, c, out, n_staged);
    }
    return hipGetLastError();
}

hipError_t launch_filter(int maxs, const ViewDesc *views, const Tile *tiles, int ntiles,
                         const PMConst &c, uint32_t *mask, hipStream_t st)
{
    if (ntiles <= 0) return hipSuccess;
    if (maxs <= 8) k_filter<8><<<ntiles, kThreads, 0, st>>>(views, tiles, c, mask);
    else if (maxs <= 16) k_filter<16><<<ntiles, kThreads, 0, st>>>(views, tiles, c, mask);
    else k_filter<32><<<ntiles, kThreads, 0, st>>>(views, tiles, c, mask);
    return hipGetLastError();
}

hipError_t launch_finalize(const ViewDesc *views, int view, int npix, float *depth, float *normal,
                           float *cost, float *depth_keep, hipStream_t st)
{
    k_finalize<<<(npix + 255) / 256, 256, 0, st>>>(views, view, depth, normal, cost, depth_keep);
    return hipGetLastError();
}

hipError_t launch_load_hyp(const ViewDesc *views, int view, int npix, const float *depth,
                           const float *normal, float cos_max_normal, hipStream_t st)
{
    k_load_hyp<<<(npix + 255) / 256, 256, 0, st>>>(views, view, depth, normal, cos_max_normal);
    return hipGetLastError();
}

hipError_t launch_advance(ViewDesc *views, int nviews, hipStream_t st)
{
    k_advance<<<(nviews + 63) / 64, 64, 0, st>>>(views, nviews);
    return hipGetLastError();
}

hipError_t launch_pack_depth4(const float *depth, int W, int H, float4 *out, hipStream_t st)
{
    const dim3 grid((W + 255) / 256, H);
    k_pack_depth4<<<grid, 256, 0, st>>>(depth, W, H, out);
    return hipGetLastError();
}

hipError_t launch_pack_pair(const uint8_t *gray, int W, int H, uint16_t *pair, hipStream_t st)
{
    dim3 grid((W + 1 + 255) / 256, H);
    k_pack_pair<<<grid, 256, 0, st>>>(gray, W, H, pair);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long *p, int n, unsigned long long v, hipStream_t st)
{
    k_fill_u64<<<(n + 255) / 256, 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

}  // namespace vsm
