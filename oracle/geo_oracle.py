"""CPU restatement of the ENU -> geodetic -> UTM chain, and the mathematics
that pins it.

TEST INFRASTRUCTURE (the oracle): imported only by tests/ and
__graft_entry__.smoke() as the checker.

* ``enu_to_utm(frame, enu)``     -- liborc_geo.so (geo_oracle.c), the chain
                                    local_to_global (coordinate_system.py:41-51)
                                    + latlon_to_eastnorh
                                    (lib/latlon_utm_converter.py:39-52) as used
                                    at aggregate_2p5d_util.py:95-98
* ``frame_constants(...)``       -- the 30 frame constants from their
                                    definitions, independently of the library:
                                    numpy trigonometry, exact rational
                                    arithmetic (fractions) for the Krueger
                                    coefficients in n = f / (2 - f)
* ``meridian_arc(phi)``          -- quadrature of the meridian radius of
                                    curvature a (1 - e^2) / (1 - e^2 sin^2)^1.5
* ``geodetic_to_enu(...)``       -- closed-form geodetic -> ECEF -> ENU (the
                                    forward direction, used to feed known
                                    geodetic points into the chain)

pymap3d / pyproj (the reference's libraries) are not installed; their exact
digits are not reproducible and this restatement is pinned by the
mathematics above instead.
"""
import ctypes
import math
import os
from fractions import Fraction

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

A_WGS84 = 6378137.0
F_WGS84 = 1.0 / 298.257223563
E2_WGS84 = F_WGS84 * (2.0 - F_WGS84)
# f = 1 / 298.257223563 exactly, so n = f / (2 - f) = 1 / (2 / f - 1)
N_EXACT = 1 / (2 * Fraction("298.257223563") - 1)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc_geo.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run make -C oracle")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_enu_to_utm.argtypes = [dp, dp, ctypes.c_int64, dp]
        L.orc_enu_to_utm.restype = None
        for fn in ("orc_atan", "orc_asinh_small", "orc_atanh_small", "orc_sinh_small"):
            getattr(L, fn).argtypes = [ctypes.c_double]
            getattr(L, fn).restype = ctypes.c_double
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def enu_to_utm(frame, enu):
    """frame: 30 FP64 constants (vsm_geo_frame order); enu: n x 3."""
    f = np.ascontiguousarray(frame, dtype=np.float64).reshape(30)
    pts = np.ascontiguousarray(enu, dtype=np.float64).reshape(-1, 3)
    out = np.zeros_like(pts)
    lib().orc_enu_to_utm(_dp(f), _dp(pts), len(pts), _dp(out))
    return out


def series(name, x):
    fn = getattr(lib(), "orc_" + name)
    return np.array([fn(float(v)) for v in np.ravel(x)])


def krueger_exact():
    """(A / a, alpha[6]) as exact rationals of n (Karney 2011, eq. 35)."""
    n = N_EXACT
    A = (1 + n ** 2 / 4 + n ** 4 / 64 + n ** 6 / 256) / (1 + n)
    F = Fraction
    alpha = [
        F(1, 2) * n - F(2, 3) * n ** 2 + F(5, 16) * n ** 3 + F(41, 180) * n ** 4
        - F(127, 288) * n ** 5 + F(7891, 37800) * n ** 6,
        F(13, 48) * n ** 2 - F(3, 5) * n ** 3 + F(557, 1440) * n ** 4 + F(281, 630) * n ** 5
        - F(1983433, 1935360) * n ** 6,
        F(61, 240) * n ** 3 - F(103, 140) * n ** 4 + F(15061, 26880) * n ** 5
        + F(167603, 181440) * n ** 6,
        F(49561, 161280) * n ** 4 - F(179, 168) * n ** 5 + F(6601661, 7257600) * n ** 6,
        F(34729, 80640) * n ** 5 - F(3418889, 1995840) * n ** 6,
        F(212378941, 319334400) * n ** 6,
    ]
    return A, alpha


def frame_constants(lat0_deg, lon0_deg, alt0, zone_number, south):
    a, f, e2 = A_WGS84, F_WGS84, E2_WGS84
    b = a * (1 - f)
    lat0, lon0 = np.radians(lat0_deg), np.radians(lon0_deg)
    sf, cf, sl, cl = np.sin(lat0), np.cos(lat0), np.sin(lon0), np.cos(lon0)
    N0 = a / np.sqrt(1 - e2 * sf * sf)
    origin = [(N0 + alt0) * cf * cl, (N0 + alt0) * cf * sl, (N0 * (1 - e2) + alt0) * sf]
    east = [-sl, cl, 0.0]
    north = [-sf * cl, -sf * sl, cf]
    up = [cf * cl, cf * sl, sf]
    rot = [east[0], north[0], up[0], east[1], north[1], up[1], east[2], north[2], up[2]]
    lam0 = np.radians(6.0 * zone_number - 183.0)
    A, alpha = krueger_exact()
    vals = origin + rot + [np.cos(lam0), np.sin(lam0), a, math.sqrt(e2), e2, e2 * a,
                           e2 / (1 - e2) * b, 1 - f, float(Fraction("0.9996") * A * Fraction(a))]
    vals += [float(x) for x in alpha]
    vals += [500000.0, 10000000.0 if south else 0.0, 0.25]
    return np.array(vals, dtype=np.float64)


def meridian_arc(phi_rad):
    """Distance along the meridian from the equator to latitude phi (m)."""
    import warnings
    from scipy.integrate import IntegrationWarning, quad
    a, e2 = A_WGS84, E2_WGS84
    f = lambda p: a * (1 - e2) / (1 - e2 * math.sin(p) ** 2) ** 1.5
    with warnings.catch_warnings():
        # the requested 1e-15 is at the FP64 limit; quad says so and still
        # returns the best estimate
        warnings.simplefilter("ignore", IntegrationWarning)
        val, _ = quad(f, 0.0, float(phi_rad), epsabs=1e-11, epsrel=1e-15, limit=200)
    return val


def geodetic_to_ecef(lat_deg, lon_deg, h):
    a, e2 = A_WGS84, E2_WGS84
    p, l = np.radians(lat_deg), np.radians(lon_deg)
    N = a / np.sqrt(1 - e2 * np.sin(p) ** 2)
    return np.stack([(N + h) * np.cos(p) * np.cos(l), (N + h) * np.cos(p) * np.sin(l),
                     (N * (1 - e2) + h) * np.sin(p)], axis=-1)


def geodetic_to_enu(lat_deg, lon_deg, h, lat0_deg, lon0_deg, alt0):
    X = geodetic_to_ecef(lat_deg, lon_deg, h)
    X0 = geodetic_to_ecef(lat0_deg, lon0_deg, alt0)
    d = X - X0
    p, l = math.radians(lat0_deg), math.radians(lon0_deg)
    east = np.array([-math.sin(l), math.cos(l), 0.0])
    north = np.array([-math.sin(p) * math.cos(l), -math.sin(p) * math.sin(l), math.cos(p)])
    up = np.array([math.cos(p) * math.cos(l), math.cos(p) * math.sin(l), math.sin(p)])
    return np.stack([d @ east, d @ north, d @ up], axis=-1)


def radii(lat_deg):
    """(meridian M, prime vertical N) radii of curvature."""
    a, e2 = A_WGS84, E2_WGS84
    s = math.sin(math.radians(lat_deg))
    w = 1 - e2 * s * s
    return a * (1 - e2) / w ** 1.5, a / math.sqrt(w)
