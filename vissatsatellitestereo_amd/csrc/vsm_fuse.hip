// vsm_fuse.hip -- 2.5D DSM fusion kernels (HBM-bound, gfx950).
//
// Per view (reference aggregate_2p5d_util.py:45-106, lib/proj_to_grid.py:41-81,
// produce_dsm.py:46-58):
//   back-projection X = inv(P4) (col, row, 1, d), divide by X3, in FP64 with a
//   fixed operation order (compiled -ffp-contract=off) so grid indices are
//   bit-exact with the CPU restatement; row = floor((ul_n - N)/e_res),
//   col = floor((E - ul_e)/n_res) (the reference's xresolution/yresolution
//   pairing, lib/proj_to_grid.py:42-43); per-cell nanmax through a 64-bit
//   atomicMax on an order-preserving encoding of the FP64 height (max is
//   order-independent, so the raster does not depend on atomic ordering);
//   NaN cells filled with the median of the in-bounds non-NaN 3x3 neighbours
//   of the unfilled raster (lib/proj_to_grid.py:64-79), cast to float32;
//   3x3 median with replicated border (produce_dsm.py:58).
// Across views (aggregate_2p5d.py:66-81): cells with fewer than
// min_measurements values -> NaN, median, MAD = median |x - med|, drop
// |x - med| > MAD, mean (FP64 accumulation), then the 3x3 median.
// Median rule (NaN-aware): NaN excluded; even count -> mean of the two middle
// values; no values -> NaN.
#include "vsm_device.h"
#include "vsm_geo.h"
#include "vsm_kernels.h"

#include <hipcub/hipcub.hpp>

#include <mutex>
#include <vector>

namespace vsm {

__device__ __forceinline__ unsigned long long enc_f64(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double dec_f64(unsigned long long k)
{
    const unsigned long long b = (k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __longlong_as_double((long long)b);
}

__device__ __forceinline__ void cell_of(const DsmGridDev &g, double E, double N, int &row, int &col,
                                        bool &ok)
{
    const double fr = floor((g.ul_n - N) / g.e_res);
    const double fc = floor((E - g.ul_e) / g.n_res);
    ok = fr >= 0.0 && fc >= 0.0 && fr < (double)g.n_size && fc < (double)g.e_size;
    row = ok ? (int)fr : 0;
    col = ok ? (int)fc : 0;
}

// One thread per depth pixel; d <= 0 or non-finite -> no point
// (aggregate_2p5d_util.py:76, 92-93).  GEO: the ENU point goes through
// local_to_global + latlon_to_eastnorh (aggregate_2p5d_util.py:95-98) and the
// cell value is the altitude.
template <bool GEO>
__global__ void k_rasterize_view(const float *__restrict__ depth, int W, int H,
                                 const double *__restrict__ Mi, DsmGridDev g, vsm_geo_frame gf,
                                 unsigned long long *__restrict__ cell_max)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= W * H) return;
    const float d32 = depth[p];
    if (!(d32 > 0.f) || !isfinite(d32)) return;
    const double col = (double)(p % W), row = (double)(p / W), d = (double)d32;
    double X[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        X[i] = ((Mi[4 * i + 0] * col + Mi[4 * i + 1] * row) + Mi[4 * i + 2]) + Mi[4 * i + 3] * d;
    double E = X[0] / X[3], N = X[1] / X[3], h = X[2] / X[3];
    if (!isfinite(E) || !isfinite(N) || !isfinite(h)) return;
    if (GEO) {
        double Eu, Nu, alt;
        if (!geo::enu_to_utm(gf, E, N, h, Eu, Nu, alt)) return;
        E = Eu;
        N = Nu;
        h = alt;
    }
    int r, c;
    bool ok;
    cell_of(g, E, N, r, c, ok);
    if (!ok) return;
    atomicMax(&cell_max[(size_t)r * g.e_size + c], enc_f64(h));
}

__global__ void k_rasterize_points(const double *__restrict__ pts, long long n, DsmGridDev g,
                                   unsigned long long *__restrict__ cell_max)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double E = pts[3 * i], N = pts[3 * i + 1], h = pts[3 * i + 2];
    if (isnan(h) || isnan(E) || isnan(N)) return;
    int r, c;
    bool ok;
    cell_of(g, E, N, r, c, ok);
    if (!ok) return;
    atomicMax(&cell_max[(size_t)r * g.e_size + c], enc_f64(h));
}

__global__ void k_enu_to_utm(const double *__restrict__ enu, long long n, vsm_geo_frame gf,
                             double *__restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double E, N, alt;
    const bool ok = geo::enu_to_utm(gf, enu[3 * i], enu[3 * i + 1], enu[3 * i + 2], E, N, alt);
    const double nan = __longlong_as_double(0x7FF8000000000000ll);
    out[3 * i] = ok ? E : nan;
    out[3 * i + 1] = ok ? N : nan;
    out[3 * i + 2] = ok ? alt : nan;
}

// ---------------------------------------------------------------------------
// 3D fusion: one thread per pixel of image i (see vsm_fuse_points).  FP64
// projections with a fixed operation order, FP32 normals.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void backproject4(const double *Mi, double col, double row, double d,
                                             double &X, double &Y, double &Z)
{
    double h[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        h[k] = ((Mi[4 * k + 0] * col + Mi[4 * k + 1] * row) + Mi[4 * k + 2]) + Mi[4 * k + 3] * d;
    X = h[0] / h[3];
    Y = h[1] / h[3];
    Z = h[2] / h[3];
}

__device__ __forceinline__ void project3(const double *P, double X, double Y, double Z, double &u0,
                                         double &u1, double &u2)
{
    u0 = ((P[0] * X + P[1] * Y) + P[2] * Z) + P[3];
    u1 = ((P[4] * X + P[5] * Y) + P[6] * Z) + P[7];
    u2 = ((P[8] * X + P[9] * Y) + P[10] * Z) + P[11];
}

__device__ __forceinline__ void rotate_normal(const float *R, const float *n, float &a, float &b,
                                              float &c)
{
    a = (R[0] * n[0] + R[1] * n[1]) + R[2] * n[2];
    b = (R[3] * n[0] + R[4] * n[1]) + R[5] * n[2];
    c = (R[6] * n[0] + R[7] * n[1]) + R[8] * n[2];
}

__global__ void k_fuse3d(const Fuse3dView *__restrict__ V, int i, Fuse3dConst k,
                         FusedPoint *__restrict__ out, int *__restrict__ flag)
{
    const Fuse3dView &A = V[i];
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)A.W * A.H) return;
    flag[p] = 0;
    const float d = A.depth[p];
    if (!(d > 0.f) || !isfinite(d)) return;
    const double col = (double)(p % A.W), row = (double)(p / A.W);
    double X, Y, Z;
    backproject4(A.Pinv, col, row, (double)d, X, Y, Z);
    if (!isfinite(X) || !isfinite(Y) || !isfinite(Z)) return;
    float na, nb, nc;
    rotate_normal(A.R, A.normal + 3 * p, na, nb, nc);
    int count = 1;
    double sx = X, sy = Y, sz = Z;
    float snx = na, sny = nb, snz = nc;
    uint32_t sg = A.gray ? A.gray[p] : 0u;
    for (int j = 0; j < k.n; ++j) {
        if (j == i) continue;
        const Fuse3dView &B = V[j];
        double u0, u1, u2;
        project3(B.P, X, Y, Z, u0, u1, u2);
        if (!(u2 > 0.0)) continue;
        const double cx = floor(u0 / u2 + 0.5), cy = floor(u1 / u2 + 0.5);
        if (!(cx >= 0.0 && cy >= 0.0 && cx < (double)B.W && cy < (double)B.H)) continue;
        const long long q = (long long)cy * B.W + (long long)cx;
        const float dj = B.depth[q];
        if (!(dj > 0.f) || !isfinite(dj)) continue;
        double Xj, Yj, Zj;
        backproject4(B.Pinv, cx, cy, (double)dj, Xj, Yj, Zj);
        const double ex = Xj - X, ey = Yj - Y, ez = Zj - Z;
        if (!(sqrt((ex * ex + ey * ey) + ez * ez) <= k.max_depth_err)) continue;
        double v0, v1, v2;
        project3(A.P, Xj, Yj, Zj, v0, v1, v2);
        if (!(v2 > 0.0)) continue;
        const double rx = v0 / v2 - col, ry = v1 / v2 - row;
        if (!(sqrt(rx * rx + ry * ry) <= k.max_reproj)) continue;
        float ma, mb, mc;
        rotate_normal(B.R, B.normal + 3 * q, ma, mb, mc);
        if (!((na * ma + nb * mb) + nc * mc >= k.cos_max_normal)) continue;
        if (j < i) return;                 // the cluster belongs to image j
        ++count;
        sx = sx + Xj;
        sy = sy + Yj;
        sz = sz + Zj;
        snx = snx + ma;
        sny = sny + mb;
        snz = snz + mc;
        sg += B.gray ? B.gray[q] : 0u;
    }
    if (count < k.min_num) return;
    const double fc = (double)count;
    const float l = sqrtf((snx * snx + sny * sny) + snz * snz);
    FusedPoint o;
    o.x = sx / fc;
    o.y = sy / fc;
    o.z = sz / fc;
    o.nx = l > 0.f ? snx / l : 0.f;
    o.ny = l > 0.f ? sny / l : 0.f;
    o.nz = l > 0.f ? snz / l : 0.f;
    o.gray = (sg + (uint32_t)(count / 2)) / (uint32_t)count;
    out[p] = o;
    flag[p] = 1;
}

__global__ void k_unpack_fused(const FusedPoint *__restrict__ s, long long n, double *__restrict__ xyz,
                               float *__restrict__ normal, uint8_t *__restrict__ gray)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const FusedPoint o = s[t];
    xyz[3 * t] = o.x;
    xyz[3 * t + 1] = o.y;
    xyz[3 * t + 2] = o.z;
    normal[3 * t] = o.nx;
    normal[3 * t + 1] = o.ny;
    normal[3 * t + 2] = o.nz;
    gray[t] = (uint8_t)min(o.gray, 255u);
}

// Small in-register sort (insertion) of up to 9 doubles/floats.
template <typename T, int N>
__device__ __forceinline__ void sort_small(T (&v)[N], int n)
{
#pragma unroll
    for (int i = 1; i < N; ++i) {
        if (i < n) {
            T x = v[i];
#pragma unroll
            for (int j = i - 1; j >= 0; --j) {
                if (v[j] > x) { v[j + 1] = v[j]; v[j] = x; }
            }
        }
    }
}

template <typename T, int N>
__device__ __forceinline__ T median_sorted(const T (&v)[N], int n)
{
    T lo = v[0], hi = v[0];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (i == (n - 1) / 2) lo = v[i];
        if (i == n / 2) hi = v[i];
    }
    return (n & 1) ? hi : (T)0.5 * (lo + hi);
}

// Decode + hole fill.  Output float (per-view DSM, then medianBlur) or double
// (proj_to_grid stand-alone).
template <typename OUT>
__global__ void k_decode_fill(const unsigned long long *__restrict__ cell_max, int e_size,
                              int n_size, OUT *__restrict__ dsm)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)e_size * n_size) return;
    const int r = (int)(i / e_size), c = (int)(i - (long long)r * e_size);
    const unsigned long long k = cell_max[i];
    if (k != 0ull) { dsm[i] = (OUT)dec_f64(k); return; }
    double v[9];
    int n = 0;
#pragma unroll
    for (int dr = -1; dr <= 1; ++dr)
#pragma unroll
        for (int dc = -1; dc <= 1; ++dc) {
            const int rr = r + dr, cc = c + dc;
            const int slot = (dr + 1) * 3 + (dc + 1);
            unsigned long long kk = 0ull;
            if (rr >= 0 && cc >= 0 && rr < n_size && cc < e_size) kk = cell_max[(size_t)rr * e_size + cc];
            v[slot] = kk ? dec_f64(kk) : (double)INFINITY;
            n += kk ? 1 : 0;
        }
    if (n == 0) { dsm[i] = (OUT)NAN; return; }
    sort_small<double, 9>(v, 9);
    dsm[i] = (OUT)median_sorted<double, 9>(v, n);
}

// 3x3 median, replicated border, NaN excluded, all-NaN -> NaN.
__global__ void k_median3(const float *__restrict__ in, int e_size, int n_size,
                          float *__restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)e_size * n_size) return;
    const int r = (int)(i / e_size), c = (int)(i - (long long)r * e_size);
    float v[9];
    int n = 0;
#pragma unroll
    for (int dr = -1; dr <= 1; ++dr)
#pragma unroll
        for (int dc = -1; dc <= 1; ++dc) {
            const int rr = min(max(r + dr, 0), n_size - 1), cc = min(max(c + dc, 0), e_size - 1);
            const float x = in[(size_t)rr * e_size + cc];
            const int slot = (dr + 1) * 3 + (dc + 1);
            v[slot] = x == x ? x : INFINITY;
            n += x == x ? 1 : 0;
        }
    if (n == 0) { out[i] = NAN; return; }
    sort_small<float, 9>(v, 9);
    out[i] = median_sorted<float, 9>(v, n);
}

// Per-view DSM in one pass over a tile of kVT_W x kVT_H cells: decode the
// cell keys, fill each empty cell with the median of its non-empty 3x3
// neighbours (the unfilled keys, as k_decode_fill), then the 3x3 median of
// the filled values (replicated border, NaN excluded, as k_median3) into
// `out`.  The keys of the tile and a 2-cell halo are staged in LDS, the
// filled values of the tile and a 1-cell halo (at border-clamped positions)
// too, so each key is read once per tile instead of by three kernels.  The
// tile's cells of `clear` (the other key grid, which the next view
// rasterises into) are zeroed; `keys` and `clear` must differ, since other
// tiles read this tile's keys as their halo.
constexpr int kVT_W = 64, kVT_H = 16;

__global__ void __launch_bounds__(256) k_view_dsm(const unsigned long long *__restrict__ keys,
                                                  unsigned long long *__restrict__ clear, int e_size,
                                                  int n_size, float *__restrict__ out)
{
    constexpr int KW = kVT_W + 4, KH = kVT_H + 4, FW = kVT_W + 2, FH = kVT_H + 2;
    __shared__ unsigned long long sk[KH * KW];
    __shared__ float sf[FH * FW];
    const int c0 = blockIdx.x * kVT_W, r0 = blockIdx.y * kVT_H;
    for (int t = threadIdx.x; t < KH * KW; t += blockDim.x) {
        const int ky = t / KW, kx = t - ky * KW;
        const int r = r0 - 2 + ky, c = c0 - 2 + kx;
        sk[t] = (r >= 0 && c >= 0 && r < n_size && c < e_size) ? keys[(size_t)r * e_size + c] : 0ull;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < FH * FW; t += blockDim.x) {
        const int fy = t / FW, fx = t - fy * FW;
        // the filled value at the border-clamped position (the median below
        // reads replicated borders)
        const int rc = min(max(r0 - 1 + fy, 0), n_size - 1), cc = min(max(c0 - 1 + fx, 0), e_size - 1);
        const int ky = rc - (r0 - 2), kx = cc - (c0 - 2);
        const unsigned long long k = sk[ky * KW + kx];
        float f;
        if (k != 0ull) {
            f = (float)dec_f64(k);
        } else {
            double v[9];
            int n = 0;
#pragma unroll
            for (int dr = -1; dr <= 1; ++dr)
#pragma unroll
                for (int dc = -1; dc <= 1; ++dc) {
                    const unsigned long long kk = sk[(ky + dr) * KW + kx + dc];
                    const int slot = (dr + 1) * 3 + (dc + 1);
                    v[slot] = kk ? dec_f64(kk) : (double)INFINITY;
                    n += kk ? 1 : 0;
                }
            if (n == 0) {
                f = NAN;
            } else {
                sort_small<double, 9>(v, 9);
                f = (float)median_sorted<double, 9>(v, n);
            }
        }
        sf[t] = f;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kVT_W * kVT_H; t += blockDim.x) {
        const int ty = t / kVT_W, tx = t - ty * kVT_W;
        const int r = r0 + ty, c = c0 + tx;
        if (r >= n_size || c >= e_size) continue;
        float v[9];
        int n = 0;
#pragma unroll
        for (int dr = -1; dr <= 1; ++dr)
#pragma unroll
            for (int dc = -1; dc <= 1; ++dc) {
                const float x = sf[(ty + 1 + dr) * FW + tx + 1 + dc];
                const int slot = (dr + 1) * 3 + (dc + 1);
                v[slot] = x == x ? x : INFINITY;
                n += x == x ? 1 : 0;
            }
        const size_t i = (size_t)r * e_size + c;
        clear[i] = 0ull;
        if (n == 0) { out[i] = NAN; continue; }
        sort_small<float, 9>(v, 9);
        out[i] = median_sorted<float, 9>(v, n);
    }
}

// Bitonic sort network over N registers (static indices only).
template <int N>
__device__ __forceinline__ void bitonic(float (&v)[N])
{
#pragma unroll
    for (int k = 2; k <= N; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    // values are finite or +inf (no NaN): min/max order them
                    const float a = v[i], b = v[l];
                    const float lo = fminf(a, b), hi = fmaxf(a, b);
                    v[i] = (i & k) == 0 ? lo : hi;
                    v[l] = (i & k) == 0 ? hi : lo;
                }
            }
}

template <int N>
__global__ void k_fuse_stack(const float *__restrict__ stack, int nviews, long long ncell,
                             int min_measurements, float *__restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncell) return;
    float v[N];
    int n = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        float x = k < nviews ? stack[(size_t)k * ncell + i] : NAN;
        const bool ok = x == x;
        n += ok ? 1 : 0;
        v[k] = ok ? x : INFINITY;
    }
    if (n < min_measurements || n == 0) { out[i] = NAN; return; }
    float s[N];
#pragma unroll
    for (int k = 0; k < N; ++k) s[k] = v[k];
    bitonic<N>(s);
    const float med = median_sorted<float, N>(s, n);
#pragma unroll
    for (int k = 0; k < N; ++k) s[k] = v[k] == INFINITY ? INFINITY : fabsf(v[k] - med);
    bitonic<N>(s);
    const float mad = median_sorted<float, N>(s, n);
    double acc = 0.0;
    int m = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (v[k] != INFINITY && !(fabsf(v[k] - med) > mad)) { acc += (double)v[k]; ++m; }
    }
    out[i] = m > 0 ? (float)(acc / (double)m) : NAN;
}

__global__ void k_fill_u32(unsigned int *p, long long n, unsigned int v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u64n(unsigned long long *p, long long n, unsigned long long v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

static inline unsigned int nblk(long long n, int b) { return (unsigned int)((n + b - 1) / b); }

hipError_t launch_rasterize_view(const float *depth, int W, int H, const double *invP4_dev,
                                 DsmGridDev g, const vsm_geo_frame *geo,
                                 unsigned long long *cell_max, hipStream_t st)
{
    if (geo)
        k_rasterize_view<true><<<nblk((long long)W * H, 256), 256, 0, st>>>(depth, W, H, invP4_dev,
                                                                            g, *geo, cell_max);
    else
        k_rasterize_view<false><<<nblk((long long)W * H, 256), 256, 0, st>>>(
            depth, W, H, invP4_dev, g, vsm_geo_frame{}, cell_max);
    return hipGetLastError();
}

hipError_t launch_fuse3d(const Fuse3dView *views, int i, int W, int H, Fuse3dConst k, FusedPoint *out,
                         int *flag, hipStream_t st)
{
    const long long n = (long long)W * H;
    if (n <= 0) return hipSuccess;
    k_fuse3d<<<nblk(n, 256), 256, 0, st>>>(views, i, k, out, flag);
    return hipGetLastError();
}

hipError_t compact_fused(const FusedPoint *pts, const int *flag, long long n, void *temp,
                         size_t &temp_bytes, FusedPoint *sel, int *n_sel, hipStream_t st)
{
    return hipcub::DeviceSelect::Flagged(temp, temp_bytes, pts, flag, sel, n_sel, (int)n, st);
}

hipError_t launch_unpack_fused(const FusedPoint *sel, long long n, double *xyz, float *normal,
                               uint8_t *gray, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_unpack_fused<<<nblk(n, 256), 256, 0, st>>>(sel, n, xyz, normal, gray);
    return hipGetLastError();
}

hipError_t launch_enu_to_utm(const double *enu, long long n, const vsm_geo_frame &geo, double *out,
                             hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_enu_to_utm<<<nblk(n, 256), 256, 0, st>>>(enu, n, geo, out);
    return hipGetLastError();
}

hipError_t launch_rasterize_points(const double *pts, long long n, DsmGridDev g,
                                   unsigned long long *cell_max, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_rasterize_points<<<nblk(n, 256), 256, 0, st>>>(pts, n, g, cell_max);
    return hipGetLastError();
}

hipError_t launch_decode_fill_f32(const unsigned long long *cell_max, int e_size, int n_size,
                                  float *dsm, hipStream_t st)
{
    k_decode_fill<float><<<nblk((long long)e_size * n_size, 256), 256, 0, st>>>(cell_max, e_size,
                                                                             n_size, dsm);
    return hipGetLastError();
}

hipError_t launch_decode_fill_f64(const unsigned long long *cell_max, int e_size, int n_size,
                                  double *dsm, hipStream_t st)
{
    k_decode_fill<double><<<nblk((long long)e_size * n_size, 256), 256, 0, st>>>(cell_max, e_size,
                                                                              n_size, dsm);
    return hipGetLastError();
}

hipError_t launch_view_dsm(const unsigned long long *keys, unsigned long long *clear, int e_size, int n_size,
                           float *out, hipStream_t st)
{
    if (keys == clear) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((e_size + kVT_W - 1) / kVT_W), (unsigned)((n_size + kVT_H - 1) / kVT_H));
    k_view_dsm<<<grid, 256, 0, st>>>(keys, clear, e_size, n_size, out);
    return hipGetLastError();
}

hipError_t launch_median3(const float *in, int e_size, int n_size, float *out, hipStream_t st)
{
    k_median3<<<nblk((long long)e_size * n_size, 256), 256, 0, st>>>(in, e_size, n_size, out);
    return hipGetLastError();
}

hipError_t launch_fuse_stack(const float *stack, int nviews, long long ncell, int min_measurements,
                             float *out, hipStream_t st)
{
    const unsigned int b = nblk(ncell, 256);
    if (nviews <= 8) k_fuse_stack<8><<<b, 256, 0, st>>>(stack, nviews, ncell, min_measurements, out);
    else if (nviews <= 16) k_fuse_stack<16><<<b, 256, 0, st>>>(stack, nviews, ncell, min_measurements, out);
    else if (nviews <= 32) k_fuse_stack<32><<<b, 256, 0, st>>>(stack, nviews, ncell, min_measurements, out);
    else if (nviews <= 64) k_fuse_stack<64><<<b, 256, 0, st>>>(stack, nviews, ncell, min_measurements, out);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned int *p, long long n, unsigned int v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_fill_u32<<<nblk(n, 256), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_u64n(unsigned long long *p, long long n, unsigned long long v,
                            hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_fill_u64n<<<nblk(n, 256), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Allocation-free rasteriser.  Its only scratch, the per-cell 64-bit max keys,
// is this file-scope device array (each device holds its own copy for the
// life of the process).  Host code never passes the symbol itself to a kernel:
// it uses the device address taken once per device by scratch_init.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned long long scratch[kNoAllocMaxCells];

static std::mutex g_scratch_mu;
static std::vector<unsigned long long *> g_scratch_ptr;   // per device ordinal

hipError_t scratch_init(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;      // the caller makes `device` current
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<unsigned long long *>(p);
    return hipSuccess;
}

hipError_t proj_to_grid_noalloc(int device, const double *pts, long long n, DsmGridDev g,
                                double *dsm, hipStream_t st)
{
    unsigned long long *cells = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (device >= 0 && device < (int)g_scratch_ptr.size()) cells = g_scratch_ptr[device];
    }
    if (!cells) return hipErrorNotInitialized;
    const long long ncell = (long long)g.e_size * g.n_size;
    if (ncell <= 0 || ncell > kNoAllocMaxCells) return hipErrorInvalidValue;
    // the previous call's keys are still there: clear this call's cells
    hipError_t e = launch_fill_u64n(cells, ncell, 0ull, st);
    if (e == hipSuccess) e = launch_rasterize_points(pts, n, g, cells, st);
    if (e == hipSuccess) e = launch_decode_fill_f64(cells, g.e_size, g.n_size, dsm, st);
    return e;
}

}  // namespace vsm
