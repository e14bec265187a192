"""Pins of the ENU -> geodetic -> UTM restatement (oracle/geo_oracle.c) and of
the host-side geodesy (vissatsatellitestereo_amd/geodesy.py), CPU only.

The reference's libraries for this step (pymap3d.enu2geodetic, pyproj 'utm',
aggregate_2p5d_util.py:95-98) are not installed, so the oracle is pinned by
the mathematics a mistake would break:
* each series against libm over its domain;
* the frame constants against an independent derivation (exact rationals for
  the Krueger coefficients);
* geodetic -> ENU (closed form) -> chain: the altitude comes back, and on the
  central meridian the easting is 500 km and the northing k0 times the
  meridian arc computed by quadrature (pins the rectifying radius and the
  Krueger coefficients);
* off the meridian the map is conformal (equal scale along meridian and
  parallel, orthogonal images), symmetric about the central meridian and maps
  the equator to northing 0;
* points outside the series' domain are reported (NaN).
"""
import math

import numpy as np
import pytest

import geo_oracle as go
from vissatsatellitestereo_amd import geodesy

FRAME = (30.3, -81.2, 0.0, 17, 0)          # ENU origin, UTM zone 17N
LAM0 = 6 * 17 - 183


@pytest.fixture(scope="module")
def frame():
    return go.frame_constants(*FRAME)


def _utm(frame, lat, lon, h=0.0):
    lat = np.atleast_1d(np.asarray(lat, dtype=np.float64))
    lon = np.atleast_1d(np.asarray(lon, dtype=np.float64))
    enu = go.geodetic_to_enu(lat, lon, np.broadcast_to(h, lat.shape), *FRAME[:3])
    return go.enu_to_utm(frame, enu)


def test_series_match_libm():
    t = np.concatenate([np.linspace(-3, 3, 2002), np.logspace(-8, 8, 400), -np.logspace(-8, 8, 400)])
    assert np.max(np.abs(go.series("atan", t) - np.arctan(t)) / np.abs(np.arctan(t))) < 1e-15
    for name, dom, ref in [("asinh_small", 0.25, np.arcsinh), ("atanh_small", 0.082, np.arctanh),
                           ("sinh_small", 0.0068, np.sinh)]:
        x = np.linspace(-dom, dom, 2002)
        err = np.abs(go.series(name, x) - ref(x)) / np.abs(ref(x))
        assert err.max() < 1e-15, name


def test_library_frame_constants_match_definitions():
    from vissatsatellitestereo_amd import _lib
    for params in [FRAME, (-33.9, 151.2, 40.0, 56, 1), (61.0, 5.0, -20.0, 32, 0), (0.2, 0.1, 0.0, 31, 0)]:
        lib = _lib.geo_frame(*params).as_array()
        ref = go.frame_constants(*params)
        scale = np.maximum(np.abs(ref), 1e-300)
        assert np.all(np.abs(lib - ref) <= 4e-16 * scale + 1e-300), params
    with pytest.raises(ValueError):
        _lib.geo_frame(30.0, 0.0, 0.0, 61, 0)
    with pytest.raises(ValueError):
        _lib.geo_frame(85.0, 0.0, 0.0, 31, 0)


def test_altitude_round_trip(frame):
    rng = np.random.default_rng(1)
    lat = np.concatenate([rng.uniform(30.2, 30.4, 500), rng.uniform(-80, 84, 500)])
    lon = np.concatenate([rng.uniform(-81.3, -81.1, 500), LAM0 + rng.uniform(-9, 9, 500)])
    h = np.concatenate([rng.uniform(-50, 200, 500), rng.uniform(-500, 9000, 500)])
    out = _utm(frame, lat, lon, h)
    assert np.max(np.abs(out[:, 2] - h)) < 1e-6


@pytest.mark.parametrize("lat", [-80.0, -45.0, -10.0, 0.0, 1e-3, 10.0, 30.3, 45.0, 60.0, 84.0])
def test_central_meridian_is_scaled_meridian_arc(frame, lat):
    out = _utm(frame, lat, LAM0, 100.0)[0]
    assert abs(out[0] - 500000.0) < 1e-8
    assert abs(out[1] - 0.9996 * go.meridian_arc(math.radians(lat))) < 2e-8


def test_conformal_off_meridian(frame):
    hdeg = 2e-5
    worst = 0.0
    for lat in [-70.0, -30.0, 0.5, 15.0, 30.3, 50.0, 75.0]:
        for dl in [-9.0, -3.0, -0.5, 0.7, 2.9, 6.0, 12.0]:
            lon = LAM0 + dl
            P = _utm(frame, [lat + hdeg, lat - hdeg, lat, lat], [lon, lon, lon + hdeg, lon - hdeg])
            dphi = (P[0, :2] - P[1, :2]) / (2 * math.radians(hdeg))
            dlam = (P[2, :2] - P[3, :2]) / (2 * math.radians(hdeg))
            M, N = go.radii(lat)
            km = np.hypot(*dphi) / M
            kp = np.hypot(*dlam) / (N * math.cos(math.radians(lat)))
            cosang = np.dot(dphi, dlam) / (np.hypot(*dphi) * np.hypot(*dlam))
            worst = max(worst, abs(km / kp - 1.0), abs(cosang))
            assert 0.9996 - 1e-9 <= km < 1.03
    assert worst < 1e-8


def test_symmetry_equator_and_south(frame):
    for dl in [0.3, 2.0, 5.0, 11.0]:
        a = _utm(frame, [0.0, 0.0, 40.0, 40.0], [LAM0 + dl, LAM0 - dl, LAM0 + dl, LAM0 - dl])
        assert abs(a[0, 1]) < 1e-8 and abs(a[1, 1]) < 1e-8
        assert abs(a[0, 0] + a[1, 0] - 1e6) < 1e-8 and abs(a[2, 0] + a[3, 0] - 1e6) < 1e-8
        assert abs(a[2, 1] - a[3, 1]) < 1e-8
        assert a[0, 0] > 500000.0
    fs = go.frame_constants(FRAME[0], FRAME[1], FRAME[2], 17, 1)
    n_north = _utm(frame, -20.0, LAM0 + 1.0)[0, 1]
    n_south = _utm(fs, -20.0, LAM0 + 1.0)[0, 1]
    assert n_south - n_north == 1e7 and 0 < n_south < 1e7


def test_domain_limit(frame):
    out = _utm(frame, [10.0, 10.0, 10.0], [LAM0 + 14.9, LAM0 - 14.9, LAM0 + 13.0])
    assert np.isnan(out[:2]).all() and np.isfinite(out[2]).all()


def test_host_utm_matches_oracle_and_inverts(frame):
    rng = np.random.default_rng(0)
    lat = rng.uniform(-80, 84, 2000)
    lon = LAM0 + rng.uniform(-8, 8, 2000)
    E, N = geodesy.latlon_to_utm(lat, lon, 17, True)
    o = _utm(frame, lat, lon)
    assert np.max(np.abs(E - o[:, 0])) < 1e-7 and np.max(np.abs(N - o[:, 1])) < 1e-7
    la, lo = geodesy.utm_to_latlon(E, N, 17, True)
    assert np.max(np.abs(la - lat)) < 1e-11 and np.max(np.abs(lo - lon)) < 1e-11


def test_host_enu_round_trip_and_zone_rule():
    rng = np.random.default_rng(2)
    e, n, u = rng.uniform(-2000, 2000, (3, 300))
    lat, lon, h = geodesy.enu_to_geodetic(e, n, u, *FRAME[:3])
    e2, n2, u2 = geodesy.geodetic_to_enu(lat, lon, h, *FRAME[:3])
    assert max(np.abs(e2 - e).max(), np.abs(n2 - n).max(), np.abs(u2 - u).max()) < 1e-6
    ref = go.geodetic_to_enu(lat, lon, h, *FRAME[:3])
    assert np.abs(ref - np.stack([e, n, u], axis=1)).max() < 1e-6
    assert geodesy.utm_zone(30.3, -81.2) == 17
    assert geodesy.utm_zone(60.0, 5.0) == 32            # Norway
    assert geodesy.utm_zone(78.0, 20.0) == 33           # Svalbard
    assert geodesy.utm_zone(-33.9, 151.2) == 56


def test_aoi_dict_and_geo_params():
    aoi = geodesy.aoi_dict(17, "N", 435000.0, 3355000.0, 400.0, 300.0, -5.0, 40.0)
    assert aoi["lr_easting"] == 435400.0 and aoi["lr_northing"] == 3354700.0
    lat, lon = geodesy.utm_to_latlon(np.array([435000.0]), np.array([3355000.0]), 17, True)
    assert aoi["lat_min"] < lat[0] <= aoi["lat_max"] and aoi["lon_min"] <= lon[0] < aoi["lon_max"]
    p = geodesy.geo_params_from_aoi(aoi)
    assert p[2:] == (-5.0, 17, 0)
    assert abs(p[0] - (aoi["lat_min"] + aoi["lat_max"]) / 2) < 1e-15
    assert geodesy.geo_params_from_aoi(dict(aoi, zone_number=0)) is None
    assert geodesy.geo_params_from_aoi(dict(aoi, frame="local ENU (east, north, up) metres")) is None
