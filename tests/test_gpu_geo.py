"""Device ENU -> geodetic -> UTM (local_to_global + latlon_to_eastnorh,
aggregate_2p5d_util.py:95-98) against the oracle (oracle/geo_oracle.c).

The device chain uses only IEEE +, -, *, / and sqrt in the oracle's order
(built with -ffp-contract=off), so every coordinate, every UTM grid index and
every DSM cell must be bit-exact; the frame constants are the library's own
(vsm_geo_frame_init), checked against their definitions on the CPU."""
import json
import os

import numpy as np
import pytest

import sites
import fusion_oracle as fo
import geo_oracle as go

pytestmark = pytest.mark.gpu
GEO = (30.3, -81.2)


@pytest.fixture(scope="module")
def ctx():
    from vissatsatellitestereo_amd.mvs import MvsContext
    c = MvsContext(device=0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def geosite(site_dir):
    return sites.build("tinygeo", sites.tiny_spec(), site_dir, geo=GEO)


def _frame(aoi):
    from vissatsatellitestereo_amd import _lib, geodesy
    return _lib.geo_frame(*geodesy.geo_params_from_aoi(aoi))


def test_enu_to_utm_bit_exact(ctx):
    from vissatsatellitestereo_amd import _lib
    rng = np.random.default_rng(7)
    for params in [(30.3, -81.2, 0.0, 17, 0), (-33.9, 151.2, 35.0, 56, 1), (69.6, 18.9, 10.0, 34, 0)]:
        f = _lib.geo_frame(*params)
        ctx.set_geo_frame(f)
        n = 200000
        enu = np.column_stack([rng.uniform(-5000, 5000, n), rng.uniform(-5000, 5000, n),
                               rng.uniform(-200, 3000, n)])
        enu[:50, 0] = np.linspace(1.0e6, 3.0e6, 50)          # beyond the series' domain
        g = ctx.enu_to_utm(enu)
        o = go.enu_to_utm(f.as_array(), enu)
        assert np.isnan(g[:50, 0]).sum() > 0
        assert np.array_equal(g, o, equal_nan=True), params
    ctx.set_geo_frame(None)


def test_enu_to_utm_without_frame_fails(ctx):
    from vissatsatellitestereo_amd._lib import VsmError
    ctx.set_geo_frame(None)
    with pytest.raises(VsmError):
        ctx.enu_to_utm(np.zeros((4, 3)))


def test_fuse_dsm_utm_bit_exact(ctx, geosite):
    aoi = json.load(open(os.path.join(geosite["wd"], "aoi.json")))
    frame = _frame(aoi)
    grid = fo.grid_from_aoi(aoi)
    rng = np.random.default_rng(12)
    ids, invs, depths = [], [], []
    for i, n in enumerate(geosite["names"]):
        d = sites.gt_depth(geosite, n).astype(np.float32)
        d = d + rng.normal(0, 0.2, d.shape).astype(np.float32)
        d[rng.random(d.shape) < 0.05] = 0.0
        h, w = d.shape
        ctx.register_geometry(300 + i, w, h, geosite["P4"][n])
        ctx.set_depth_map(300 + i, d, kind=1)
        ids.append(300 + i)
        invs.append(np.linalg.inv(geosite["P4"][n]))
        depths.append(d)
    ctx.set_geo_frame(frame)
    try:
        fused, pv = ctx.fuse_dsm(grid, ids, depth_kind=1, min_measurements=3,
                                 inv_proj_mats=np.stack(invs), per_view=True)
    finally:
        ctx.set_geo_frame(None)
    views = [fo.view_dsm(d, inv, grid, frame.as_array()) for d, inv in zip(depths, invs)]
    for k in range(len(ids)):
        assert np.array_equal(pv[k], views[k], equal_nan=True), k
    assert np.array_equal(fused, fo.fuse(np.stack(views), 3), equal_nan=True)
    gt = np.load(os.path.join(geosite["wd"], "synthetic_gt", "gt_dsm_utm.npy"))
    err, cov = fo.median_abs_error(fused, gt)
    print("UTM DSM of noisy GT depths: median |dh|", err, "coverage", cov)
    assert err < 0.3 and cov > 0.5


def test_dropin_pipeline_on_utm_site(tmp_path_factory):
    """run_photometric_mvs -> run_consistency_check -> run_fuse on a site
    whose aoi.json is in UTM: GeoTIFFs in zone 17N, fused DSM bit-exact with
    the oracle, altitude error against the UTM ground truth."""
    from vissatsatellitestereo_amd import aggregate_2p5d, runtime
    from vissatsatellitestereo_amd import colmap_mvs_commands as cmc
    from vissatsatellitestereo_amd.lib.dsm_util import read_dsm_tif
    ws = sites.build("geo_dropin", sites.tiny_spec(), tmp_path_factory.mktemp("geo"), geo=GEO)
    runtime.reset_context()
    cmc.run_photometric_mvs(ws["mvs"], 4, num_iterations=3)
    cmc.run_consistency_check(ws["mvs"], 4, num_iterations=3)
    fused = aggregate_2p5d.run_fuse(ws["wd"], min_measurements=3)
    aoi = json.load(open(os.path.join(ws["wd"], "aoi.json")))
    ofused, _ = fo.run_fuse(ws["wd"], "geometric", 3, frame=_frame(aoi).as_array())
    assert np.array_equal(fused, ofused, equal_nan=True)
    img, meta = read_dsm_tif(os.path.join(ws["wd"], "mvs_results", "aggregate_2p5d", "aggregate_2p5d_dsm.tif"))
    assert np.array_equal(img, fused, equal_nan=True)
    assert (meta["zone_number"], meta["hemisphere"]) == (17, "N")
    assert meta["ul_easting"] == aoi["ul_easting"] and meta["ul_northing"] == aoi["ul_northing"]
    gt = np.load(os.path.join(ws["wd"], "synthetic_gt", "gt_dsm_utm.npy"))
    err, cov = fo.median_abs_error(fused, gt)
    print("drop-in UTM DSM median |dh|", err, "coverage", cov)
    assert err < 0.5 and cov > 0.3
