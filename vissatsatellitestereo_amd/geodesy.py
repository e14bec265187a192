"""Host-side WGS84 geodesy: the AOI frame, UTM <-> geodetic, geodetic <-> ENU.

The per-point ENU -> UTM conversion of the fusion path runs on the device
(``vsm_enu_to_utm`` / ``vsm_set_geo_frame``); this module holds the host side
around it:

* ``geo_params_from_aoi`` -- the ENU origin of coordinate_system.py:41-51
  (lat0, lon0 = AOI centre, alt0 = alt_min) plus aoi.json's UTM zone.  An
  aoi.json whose ``zone_number`` is 0 or that carries ``"frame": "local ENU"``
  (the synthetic sites of this package) has no geodetic anchor: its DSM grid
  is in the local ENU frame.
* ``utm_zone`` -- the zone rule of ``utm.from_latlon`` used at
  lib/latlon_utm_converter.py:48 (6-degree zones, the Norway and Svalbard
  exceptions).
* ``latlon_to_utm`` / ``utm_to_latlon`` -- WGS84 transverse Mercator by
  Krueger's series to order n^6 (Karney 2011, J. Geodesy 85:475, eqs. 7-11,
  35-36), in numpy; the reference calls pyproj here
  (lib/latlon_utm_converter.py:39-64) and ``utm.to_latlon`` when it builds
  aoi.json (stereo_pipeline.py:185-226).
* ``geodetic_to_ecef``, ``ecef_to_enu``, ``geodetic_to_enu``,
  ``enu_to_geodetic`` -- pymap3d's geodetic2enu / enu2geodetic
  (lib/latlonalt_enu_converter.py:37-46).
* ``aoi_dict`` -- the aoi.json of stereo_pipeline.py:185-226.
"""
import math

import numpy as np

WGS84_A = 6378137.0
WGS84_F = 1.0 / 298.257223563
WGS84_E2 = WGS84_F * (2.0 - WGS84_F)
WGS84_B = WGS84_A * (1.0 - WGS84_F)
UTM_K0 = 0.9996
UTM_FE = 500000.0
UTM_FN_SOUTH = 10000000.0


def krueger_coefficients():
    """(A, alpha[6], beta[6]) of the order-n^6 Krueger series for WGS84."""
    n = WGS84_F / (2.0 - WGS84_F)
    n2, n3, n4, n5, n6 = n ** 2, n ** 3, n ** 4, n ** 5, n ** 6
    A = WGS84_A / (1.0 + n) * (1.0 + n2 / 4.0 + n4 / 64.0 + n6 / 256.0)
    alpha = [
        n / 2 - 2 * n2 / 3 + 5 * n3 / 16 + 41 * n4 / 180 - 127 * n5 / 288 + 7891 * n6 / 37800,
        13 * n2 / 48 - 3 * n3 / 5 + 557 * n4 / 1440 + 281 * n5 / 630 - 1983433 * n6 / 1935360,
        61 * n3 / 240 - 103 * n4 / 140 + 15061 * n5 / 26880 + 167603 * n6 / 181440,
        49561 * n4 / 161280 - 179 * n5 / 168 + 6601661 * n6 / 7257600,
        34729 * n5 / 80640 - 3418889 * n6 / 1995840,
        212378941 * n6 / 319334400,
    ]
    beta = [
        n / 2 - 2 * n2 / 3 + 37 * n3 / 96 - n4 / 360 - 81 * n5 / 512 + 96199 * n6 / 604800,
        n2 / 48 + n3 / 15 - 437 * n4 / 1440 + 46 * n5 / 105 - 1118711 * n6 / 3870720,
        17 * n3 / 480 - 37 * n4 / 840 - 209 * n5 / 4480 + 5569 * n6 / 90720,
        4397 * n4 / 161280 - 11 * n5 / 504 - 830251 * n6 / 7257600,
        4583 * n5 / 161280 - 108847 * n6 / 3991680,
        20648693 * n6 / 638668800,
    ]
    return A, np.array(alpha), np.array(beta)


def utm_zone(lat, lon):
    """Zone number for a point (lat, lon in degrees)."""
    if 56.0 <= lat < 64.0 and 3.0 <= lon < 12.0:
        return 32
    if 72.0 <= lat <= 84.0 and lon >= 0.0:
        if lon < 9.0:
            return 31
        if lon < 21.0:
            return 33
        if lon < 33.0:
            return 35
        if lon < 42.0:
            return 37
    return int((lon + 180.0) / 6.0) + 1


def central_meridian(zone):
    return 6.0 * zone - 183.0


def _tau_prime(tau):
    e = math.sqrt(WGS84_E2)
    sig = np.sinh(e * np.arctanh(e * tau / np.sqrt(1.0 + tau * tau)))
    return tau * np.sqrt(1.0 + sig * sig) - sig * np.sqrt(1.0 + tau * tau)


def latlon_to_utm(lat, lon, zone, northern=True):
    """Degrees -> (easting, northing) in metres."""
    A, alpha, _ = krueger_coefficients()
    phi = np.radians(np.asarray(lat, dtype=np.float64))
    lam = np.radians(np.asarray(lon, dtype=np.float64) - central_meridian(zone))
    taup = _tau_prime(np.tan(phi))
    xip = np.arctan2(taup, np.cos(lam))
    etap = np.arcsinh(np.sin(lam) / np.hypot(taup, np.cos(lam)))
    xi, eta = xip.copy(), etap.copy()
    for j in range(1, 7):
        xi = xi + alpha[j - 1] * np.sin(2 * j * xip) * np.cosh(2 * j * etap)
        eta = eta + alpha[j - 1] * np.cos(2 * j * xip) * np.sinh(2 * j * etap)
    E = UTM_FE + UTM_K0 * A * eta
    N = UTM_K0 * A * xi + (0.0 if northern else UTM_FN_SOUTH)
    return E, N


def utm_to_latlon(easting, northing, zone, northern=True):
    """(easting, northing) in metres -> degrees (lat, lon)."""
    A, _, beta = krueger_coefficients()
    xi = (np.asarray(northing, dtype=np.float64) - (0.0 if northern else UTM_FN_SOUTH)) / (UTM_K0 * A)
    eta = (np.asarray(easting, dtype=np.float64) - UTM_FE) / (UTM_K0 * A)
    xip, etap = xi.copy(), eta.copy()
    for j in range(1, 7):
        xip = xip - beta[j - 1] * np.sin(2 * j * xi) * np.cosh(2 * j * eta)
        etap = etap - beta[j - 1] * np.cos(2 * j * xi) * np.sinh(2 * j * eta)
    taup = np.sin(xip) / np.hypot(np.sinh(etap), np.cos(xip))
    lam = np.arctan2(np.sinh(etap), np.cos(xip))
    # Newton on tau'(tau) = taup (Karney 2011, eqs. 19-21)
    e2 = WGS84_E2
    tau = taup.copy()
    for _ in range(6):
        tp = _tau_prime(tau)
        dtau = (taup - tp) / ((1.0 - e2) * np.sqrt(1.0 + tp * tp) * np.sqrt(1.0 + tau * tau)
                              / (1.0 + (1.0 - e2) * tau * tau))
        tau = tau + dtau
    return np.degrees(np.arctan(tau)), np.degrees(lam) + central_meridian(zone)


def geodetic_to_ecef(lat, lon, alt):
    phi = np.radians(np.asarray(lat, dtype=np.float64))
    lam = np.radians(np.asarray(lon, dtype=np.float64))
    N = WGS84_A / np.sqrt(1.0 - WGS84_E2 * np.sin(phi) ** 2)
    x = (N + alt) * np.cos(phi) * np.cos(lam)
    y = (N + alt) * np.cos(phi) * np.sin(lam)
    z = (N * (1.0 - WGS84_E2) + alt) * np.sin(phi)
    return x, y, z


def _enu_basis(lat0, lon0):
    p, l = math.radians(lat0), math.radians(lon0)
    east = np.array([-math.sin(l), math.cos(l), 0.0])
    north = np.array([-math.sin(p) * math.cos(l), -math.sin(p) * math.sin(l), math.cos(p)])
    up = np.array([math.cos(p) * math.cos(l), math.cos(p) * math.sin(l), math.sin(p)])
    return east, north, up


def ecef_to_enu(x, y, z, lat0, lon0, alt0):
    x0, y0, z0 = geodetic_to_ecef(lat0, lon0, alt0)
    d = np.stack([np.asarray(x) - x0, np.asarray(y) - y0, np.asarray(z) - z0], axis=-1)
    east, north, up = _enu_basis(lat0, lon0)
    return d @ east, d @ north, d @ up


def enu_to_ecef(e, n, u, lat0, lon0, alt0):
    x0, y0, z0 = geodetic_to_ecef(lat0, lon0, alt0)
    east, north, up = _enu_basis(lat0, lon0)
    e, n, u = (np.asarray(v, dtype=np.float64) for v in (e, n, u))
    return (x0 + e * east[0] + n * north[0] + u * up[0],
            y0 + e * east[1] + n * north[1] + u * up[1],
            z0 + e * east[2] + n * north[2] + u * up[2])


def ecef_to_geodetic(x, y, z, iterations=6):
    """Latitude/longitude in degrees and height (Bowring's iteration)."""
    x, y, z = (np.asarray(v, dtype=np.float64) for v in (x, y, z))
    f = WGS84_F
    ep2 = WGS84_E2 / (1.0 - WGS84_E2)
    p = np.hypot(x, y)
    beta = np.arctan2(z, (1.0 - f) * p)
    for _ in range(iterations):
        phi = np.arctan2(z + ep2 * WGS84_B * np.sin(beta) ** 3, p - WGS84_E2 * WGS84_A * np.cos(beta) ** 3)
        beta = np.arctan2((1.0 - f) * np.sin(phi), np.cos(phi))
    h = p * np.cos(phi) + z * np.sin(phi) - WGS84_A * np.sqrt(1.0 - WGS84_E2 * np.sin(phi) ** 2)
    return np.degrees(phi), np.degrees(np.arctan2(y, x)), h


def geodetic_to_enu(lat, lon, alt, lat0, lon0, alt0):
    return ecef_to_enu(*geodetic_to_ecef(lat, lon, alt), lat0, lon0, alt0)


def enu_to_geodetic(e, n, u, lat0, lon0, alt0):
    return ecef_to_geodetic(*enu_to_ecef(e, n, u, lat0, lon0, alt0))


def aoi_dict(zone_number, hemisphere, ul_easting, ul_northing, width, height, alt_min, alt_max):
    """aoi.json of stereo_pipeline.py:185-226 (lat/lon box of the 4 UTM corners)."""
    lr_e, lr_n = ul_easting + width, ul_northing - height
    ce = [ul_easting, lr_e, lr_e, ul_easting]
    cn = [ul_northing, ul_northing, lr_n, lr_n]
    lat, lon = utm_to_latlon(np.array(ce), np.array(cn), zone_number, hemisphere == "N")
    return {"zone_number": int(zone_number), "hemisphere": hemisphere,
            "ul_easting": float(ul_easting), "ul_northing": float(ul_northing),
            "lr_easting": float(lr_e), "lr_northing": float(lr_n),
            "width": float(width), "height": float(height),
            "lat_min": float(lat.min()), "lat_max": float(lat.max()),
            "lon_min": float(lon.min()), "lon_max": float(lon.max()),
            "alt_min": float(alt_min), "alt_max": float(alt_max)}


def geo_params_from_aoi(aoi):
    """(lat0, lon0, alt0, zone_number, south) or None for a local-ENU aoi.json."""
    if int(aoi.get("zone_number", 0)) == 0 or str(aoi.get("frame", "")).startswith("local"):
        return None
    lat0 = (aoi["lat_min"] + aoi["lat_max"]) / 2.0
    lon0 = (aoi["lon_min"] + aoi["lon_max"]) / 2.0
    return (lat0, lon0, float(aoi["alt_min"]), int(aoi["zone_number"]),
            0 if aoi["hemisphere"] == "N" else 1)
