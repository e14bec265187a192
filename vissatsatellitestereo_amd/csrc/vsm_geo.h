// vsm_geo.h -- local ENU -> WGS84 geodetic -> UTM (E, N, altitude) in FP64 on
// the device.
//
// Reference chain (aggregate_2p5d_util.py:95-98):
//   local_to_global (coordinate_system.py:41-51): pymap3d.enu2geodetic about
//     the AOI centre (lat0, lon0) at alt0 = alt_min
//     (lib/latlonalt_enu_converter.py:43-46);
//   latlon_to_eastnorh (lib/latlon_utm_converter.py:39-52): pyproj 'utm' on
//     WGS84 in the zone / hemisphere of the AOI.
// Method (own arithmetic, only IEEE +, -, *, / and sqrt, so the compiled code
// with -ffp-contract=off gives the same bits on the CPU and the GPU):
//   ENU -> ECEF: origin + R (e, n, u), R = the ENU basis in ECEF (the rotation
//     pymap3d's enu2uvw applies);
//   ECEF -> geodetic latitude and height: Bowring's iteration on the
//     parametric latitude (Bowring 1976, Survey Review 23:181), a fixed three
//     rounds, with angles carried as normalised (cos, sin) pairs;
//     h = p cos(phi) + Z sin(phi) - a sqrt(1 - e^2 sin^2(phi));
//   geodetic -> UTM: Krueger's series in the third flattening n to order n^6
//     as given by Karney (2011, J. Geodesy 85:475, eqs. 7-11 and 35):
//     tau' = tau sqrt(1 + sigma^2) - sigma sqrt(1 + tau^2),
//     sigma = sinh(e atanh(e sin(phi))), xi' = atan2(tau', cos(dlambda)),
//     eta' = asinh(sin(dlambda) / sqrt(tau'^2 + cos^2(dlambda))),
//     xi = xi' + sum_j alpha_j sin(2j xi') cosh(2j eta'),
//     eta = eta' + sum_j alpha_j cos(2j xi') sinh(2j eta'),
//     E = FE + k0 A eta, N = FN + k0 A xi.
// Special functions are Taylor series evaluated by Horner's rule on the
// domains they are used on (atan after two argument halvings; asinh, atanh and
// sinh only for small arguments: a point whose |sin(dlambda)| /
// sqrt(tau'^2 + cos^2(dlambda)) exceeds frame.max_q, i.e. more than about 14
// degrees of longitude from the central meridian, is reported invalid).
#pragma once

#include "vissat_mvs.h"

#define VSM_GEO_FN __device__ __forceinline__

namespace vsm {
namespace geo {

VSM_GEO_FN void normalise(double x, double y, double &c, double &s)
{
    const double r = sqrt(x * x + y * y);
    c = x / r;
    s = y / r;
}

// atan(t) for |t| <= tan(pi/16): sum_k (-1)^k t^(2k+1) / (2k+1), k = 0..12.
VSM_GEO_FN double atan_small(double t)
{
    const double z = t * t;
    double p = 0.04;
    p = p * z + -0.043478260869565216;
    p = p * z + 0.047619047619047616;
    p = p * z + -0.05263157894736842;
    p = p * z + 0.058823529411764705;
    p = p * z + -0.06666666666666667;
    p = p * z + 0.07692307692307693;
    p = p * z + -0.09090909090909091;
    p = p * z + 0.1111111111111111;
    p = p * z + -0.14285714285714285;
    p = p * z + 0.2;
    p = p * z + -0.3333333333333333;
    p = p * z;
    return t + t * p;
}

// atan(t) for any finite t: |t| > 1 through pi/2 - atan(1/t), then two
// halvings atan(t) = 2 atan(t / (1 + sqrt(1 + t^2))).
VSM_GEO_FN double atan_any(double t)
{
    const double kHalfPi = 1.5707963267948966;
    const bool neg = t < 0.0;
    double a = neg ? -t : t;
    const bool inv = a > 1.0;
    if (inv) a = 1.0 / a;
    a = a / (1.0 + sqrt(1.0 + a * a));
    a = a / (1.0 + sqrt(1.0 + a * a));
    double r = 4.0 * atan_small(a);
    if (inv) r = kHalfPi - r;
    return neg ? -r : r;
}

// atan2(y, x) for x > 0.
VSM_GEO_FN double atan2_xpos(double y, double x) { return atan_any(y / x); }

// asinh(q) for |q| <= 0.25: sum_n (-1)^n (2n)! / (4^n (n!)^2 (2n+1)) q^(2n+1).
VSM_GEO_FN double asinh_small(double q)
{
    const double z = q * q;
    double p = -0.004660143486915096;
    p = p * z + 0.005153309682319905;
    p = p * z + -0.005740037670841924;
    p = p * z + 0.006447210311889649;
    p = p * z + -0.0073125258735988454;
    p = p * z + 0.008390335809616815;
    p = p * z + -0.009761609529194078;
    p = p * z + 0.011551800896139705;
    p = p * z + -0.01396484375;
    p = p * z + 0.017352764423076924;
    p = p * z + -0.022372159090909092;
    p = p * z + 0.030381944444444444;
    p = p * z + -0.044642857142857144;
    p = p * z + 0.075;
    p = p * z + -0.16666666666666666;
    p = p * z;
    return q + q * p;
}

// atanh(y) for |y| <= 0.0819 (y = e sin(phi)): sum_k y^(2k+1) / (2k+1), k = 0..8.
VSM_GEO_FN double atanh_small(double y)
{
    const double z = y * y;
    double p = 0.058823529411764705;
    p = p * z + 0.06666666666666667;
    p = p * z + 0.07692307692307693;
    p = p * z + 0.09090909090909091;
    p = p * z + 0.1111111111111111;
    p = p * z + 0.14285714285714285;
    p = p * z + 0.2;
    p = p * z + 0.3333333333333333;
    p = p * z;
    return y + y * p;
}

// sinh(x) for |x| <= 0.0068: sum_k x^(2k+1) / (2k+1)!, k = 0..4.
VSM_GEO_FN double sinh_small(double x)
{
    const double z = x * x;
    double p = 2.7557319223985893e-06;
    p = p * z + 0.0001984126984126984;
    p = p * z + 0.008333333333333333;
    p = p * z + 0.16666666666666666;
    p = p * z;
    return x + x * p;
}

// Local ENU (e, n, u) -> (E, N, altitude).  Returns false for a point outside
// the series' domain (or a degenerate one), which the caller drops.
VSM_GEO_FN bool enu_to_utm(const vsm_geo_frame &f, double e, double n, double u, double &E,
                           double &N, double &alt)
{
    const double *R = f.rot;
    const double X = f.origin[0] + ((R[0] * e + R[1] * n) + R[2] * u);
    const double Y = f.origin[1] + ((R[3] * e + R[4] * n) + R[5] * u);
    const double Z = f.origin[2] + ((R[6] * e + R[7] * n) + R[8] * u);
    // longitude relative to the central meridian, as (cos, sin)
    const double xr = X * f.cm_cos + Y * f.cm_sin;
    const double yr = Y * f.cm_cos - X * f.cm_sin;
    const double p = sqrt(X * X + Y * Y);
    if (!(p > 0.0) || !(xr > 0.0)) return false;
    double cl, sl;
    normalise(xr, yr, cl, sl);
    // Bowring: parametric latitude beta, geodetic latitude phi
    double cb, sb, cp = 1.0, sp = 0.0;
    normalise(f.one_minus_f * p, Z, cb, sb);
    for (int it = 0; it < 3; ++it) {
        normalise(p - f.e2a * ((cb * cb) * cb), Z + f.ep2b * ((sb * sb) * sb), cp, sp);
        normalise(cp, f.one_minus_f * sp, cb, sb);
    }
    alt = (p * cp + Z * sp) - f.a * sqrt(1.0 - f.e2 * (sp * sp));
    // conformal latitude (Karney's tau')
    const double tau = sp / cp;
    const double sig = sinh_small(f.e * atanh_small(f.e * sp));
    const double taup = tau * sqrt(1.0 + sig * sig) - sig * sqrt(1.0 + tau * tau);
    const double q = sl / sqrt(taup * taup + cl * cl);
    if (!(q <= f.max_q && q >= -f.max_q)) return false;
    const double xip = atan2_xpos(taup, cl);
    const double etap = asinh_small(q);
    // multiple angles: (sin, cos)(2j xi'), (sinh, cosh)(2j eta'), j = 1..6
    double cx, sx;
    normalise(cl, taup, cx, sx);
    const double s2 = 2.0 * (sx * cx), c2 = cx * cx - sx * sx;
    const double sh2 = 2.0 * (q * sqrt(1.0 + q * q)), ch2 = 1.0 + 2.0 * (q * q);
    double sj[6], cj[6], shj[6], chj[6];
    sj[0] = s2; cj[0] = c2; shj[0] = sh2; chj[0] = ch2;
#pragma unroll
    for (int j = 1; j < 6; ++j) {
        sj[j] = sj[j - 1] * c2 + cj[j - 1] * s2;
        cj[j] = cj[j - 1] * c2 - sj[j - 1] * s2;
        shj[j] = shj[j - 1] * ch2 + chj[j - 1] * sh2;
        chj[j] = chj[j - 1] * ch2 + shj[j - 1] * sh2;
    }
    double dxi = 0.0, deta = 0.0;
#pragma unroll
    for (int j = 5; j >= 0; --j) {
        dxi = dxi + f.alpha[j] * (sj[j] * chj[j]);
        deta = deta + f.alpha[j] * (cj[j] * shj[j]);
    }
    N = f.false_northing + f.k0A * (xip + dxi);
    E = f.false_easting + f.k0A * (etap + deta);
    return true;
}

}  // namespace geo
}  // namespace vsm
