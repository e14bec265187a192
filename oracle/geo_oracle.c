/* geo_oracle.c -- CPU restatement of the reference's ENU -> geodetic -> UTM
 * chain (aggregate_2p5d_util.py:95-98).
 *
 * TEST INFRASTRUCTURE (the oracle): loaded only by tests/ and
 * __graft_entry__.smoke() as the checker.
 *
 *   local_to_global     coordinate_system.py:41-51 -> pymap3d.enu2geodetic
 *                       (lib/latlonalt_enu_converter.py:43-46): ENU -> ECEF
 *                       about the origin, ECEF -> geodetic
 *   latlon_to_eastnorh  lib/latlon_utm_converter.py:39-52 -> pyproj 'utm',
 *                       WGS84
 *
 * pymap3d 1.7.15 and pyproj 2.4.0 (requirements.txt:5-16) are not installed,
 * so their exact digits are not reproducible here.  This restatement uses
 * published closed forms (Bowring's parametric-latitude iteration for ECEF ->
 * geodetic, Krueger's n-series to order n^6 for the transverse Mercator in the
 * form of Karney 2011) and is pinned by the tests against the mathematics:
 * libm for every series, geodetic -> ECEF -> geodetic round trips, the
 * meridian arc by quadrature on the central meridian, conformality (equal
 * scale along the meridian and the parallel, orthogonal images) off it.
 * The arithmetic is plain IEEE double (+, -, *, /, sqrt; built with
 * -ffp-contract=off) in a fixed order; the frame constants come in as the
 * 30 doubles of vsm_geo_frame (origin[3], rot[9], cm_cos, cm_sin, a, e, e2,
 * e2a, ep2b, one_minus_f, k0A, alpha[6], false_easting, false_northing,
 * max_q), checked separately against their definitions.
 */
#include <math.h>
#include <stdint.h>

enum {
    F_ORIGIN = 0, F_ROT = 3, F_CMC = 12, F_CMS = 13, F_A = 14, F_E = 15, F_E2 = 16, F_E2A = 17,
    F_EP2B = 18, F_OMF = 19, F_K0A = 20, F_ALPHA = 21, F_FE = 27, F_FN = 28, F_MAXQ = 29,
    F_COUNT = 30
};

static void unit2(double x, double y, double *c, double *s)
{
    double r = sqrt(x * x + y * y);
    *c = x / r;
    *s = y / r;
}

/* Horner evaluation of sum_k c[k] z^k for k = 1..m-1 (c[0] = 1 is implicit):
 * x + x * (z (c1 + z (c2 + ...))) */
static double odd_series(double x, const double *c, int m)
{
    double z = x * x;
    double p = c[m - 1];
    for (int k = m - 2; k >= 1; --k) p = p * z + c[k];
    p = p * z;
    return x + x * p;
}

/* (-1)^k / (2k+1) */
static const double ATAN_C[13] = {1.0, -0.3333333333333333, 0.2, -0.14285714285714285,
                                  0.1111111111111111, -0.09090909090909091, 0.07692307692307693,
                                  -0.06666666666666667, 0.058823529411764705, -0.05263157894736842,
                                  0.047619047619047616, -0.043478260869565216, 0.04};
/* (-1)^n (2n)! / (4^n (n!)^2 (2n+1)) */
static const double ASINH_C[16] = {
    1.0, -0.16666666666666666, 0.075, -0.044642857142857144, 0.030381944444444444,
    -0.022372159090909092, 0.017352764423076924, -0.01396484375, 0.011551800896139705,
    -0.009761609529194078, 0.008390335809616815, -0.0073125258735988454, 0.006447210311889649,
    -0.005740037670841924, 0.005153309682319905, -0.004660143486915096};
/* 1 / (2k+1) */
static const double ATANH_C[9] = {1.0, 0.3333333333333333, 0.2, 0.14285714285714285,
                                  0.1111111111111111, 0.09090909090909091, 0.07692307692307693,
                                  0.06666666666666667, 0.058823529411764705};
/* 1 / (2k+1)! */
static const double SINH_C[5] = {1.0, 0.16666666666666666, 0.008333333333333333,
                                 0.0001984126984126984, 2.7557319223985893e-06};

double orc_atan(double t)
{
    int neg = t < 0.0;
    double a = neg ? -t : t;
    int inv = a > 1.0;
    if (inv) a = 1.0 / a;
    for (int h = 0; h < 2; ++h) a = a / (1.0 + sqrt(1.0 + a * a));
    double r = 4.0 * odd_series(a, ATAN_C, 13);
    if (inv) r = 1.5707963267948966 - r;
    return neg ? -r : r;
}
double orc_asinh_small(double q) { return odd_series(q, ASINH_C, 16); }
double orc_atanh_small(double y) { return odd_series(y, ATANH_C, 9); }
double orc_sinh_small(double x) { return odd_series(x, SINH_C, 5); }

/* One point; returns 0 when outside the series' domain. */
static int enu_to_utm1(const double *F, double e, double n, double u, double *out)
{
    const double *R = F + F_ROT;
    double X = F[F_ORIGIN + 0] + ((R[0] * e + R[1] * n) + R[2] * u);
    double Y = F[F_ORIGIN + 1] + ((R[3] * e + R[4] * n) + R[5] * u);
    double Z = F[F_ORIGIN + 2] + ((R[6] * e + R[7] * n) + R[8] * u);
    double xr = X * F[F_CMC] + Y * F[F_CMS];
    double yr = Y * F[F_CMC] - X * F[F_CMS];
    double p = sqrt(X * X + Y * Y);
    if (!(p > 0.0) || !(xr > 0.0)) return 0;
    double cl, sl, cb, sb, cp = 1.0, sp = 0.0;
    unit2(xr, yr, &cl, &sl);
    unit2(F[F_OMF] * p, Z, &cb, &sb);
    for (int it = 0; it < 3; ++it) {
        unit2(p - F[F_E2A] * ((cb * cb) * cb), Z + F[F_EP2B] * ((sb * sb) * sb), &cp, &sp);
        unit2(cp, F[F_OMF] * sp, &cb, &sb);
    }
    double alt = (p * cp + Z * sp) - F[F_A] * sqrt(1.0 - F[F_E2] * (sp * sp));
    double tau = sp / cp;
    double sig = orc_sinh_small(F[F_E] * orc_atanh_small(F[F_E] * sp));
    double taup = tau * sqrt(1.0 + sig * sig) - sig * sqrt(1.0 + tau * tau);
    double q = sl / sqrt(taup * taup + cl * cl);
    if (!(q <= F[F_MAXQ] && q >= -F[F_MAXQ])) return 0;
    double xip = orc_atan(taup / cl);
    double etap = orc_asinh_small(q);
    double cx, sx;
    unit2(cl, taup, &cx, &sx);
    double s2 = 2.0 * (sx * cx), c2 = cx * cx - sx * sx;
    double sh2 = 2.0 * (q * sqrt(1.0 + q * q)), ch2 = 1.0 + 2.0 * (q * q);
    double sj[6], cj[6], shj[6], chj[6];
    sj[0] = s2; cj[0] = c2; shj[0] = sh2; chj[0] = ch2;
    for (int j = 1; j < 6; ++j) {
        sj[j] = sj[j - 1] * c2 + cj[j - 1] * s2;
        cj[j] = cj[j - 1] * c2 - sj[j - 1] * s2;
        shj[j] = shj[j - 1] * ch2 + chj[j - 1] * sh2;
        chj[j] = chj[j - 1] * ch2 + shj[j - 1] * sh2;
    }
    double dxi = 0.0, deta = 0.0;
    for (int j = 5; j >= 0; --j) {
        dxi = dxi + F[F_ALPHA + j] * (sj[j] * chj[j]);
        deta = deta + F[F_ALPHA + j] * (cj[j] * shj[j]);
    }
    out[0] = F[F_FE] + F[F_K0A] * (etap + deta);
    out[1] = F[F_FN] + F[F_K0A] * (xip + dxi);
    out[2] = alt;
    return 1;
}

/* n x 3 ENU -> n x 3 (E, N, altitude); NaN rows outside the domain. */
void orc_enu_to_utm(const double *frame, const double *enu, int64_t n, double *out)
{
    for (int64_t i = 0; i < n; ++i) {
        if (!enu_to_utm1(frame, enu[3 * i], enu[3 * i + 1], enu[3 * i + 2], out + 3 * i)) {
            out[3 * i] = NAN;
            out[3 * i + 1] = NAN;
            out[3 * i + 2] = NAN;
        }
    }
}
