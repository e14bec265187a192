"""Device 3D fusion (vsm_fuse_points) against the oracle
(oracle/fusion3d_oracle.py): FP64 projections and FP32 normals in the same
operation order, so the point set, its order and every coordinate, normal and
grey value must be identical; then the aggregate_3d drop-in end to end."""
import json
import os

import numpy as np
import pytest

import sites
import fusion3d_oracle as f3
from test_oracle_fusion3d import _inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tiny(site_dir):
    return sites.build("tiny", sites.tiny_spec(), site_dir)


def test_fuse_points_bit_exact(tiny):
    from vissatsatellitestereo_amd.mvs import MvsContext
    ctx = MvsContext(device=0)
    try:
        tilt = dict(min_num_pixels=3, max_reproj_error=1.0)
        for noise, kw in [(0.0, {}), (0.3, {}), (0.3, dict(min_num_pixels=2, max_depth_error=0.5)),
                          (0.2, dict(tilt, max_normal_error_deg=40.0)),
                          (0.2, dict(tilt, max_normal_error_deg=10.0))]:
            depths, normals, grays, P4s, invs, rots = _inputs(tiny, noise=noise, seed=5)
            if "max_normal_error_deg" in kw:
                # image 1's normals turned by 20 degrees about its camera x axis
                c, s_ = np.cos(np.radians(20.0)), np.sin(np.radians(20.0))
                Rx = np.array([[1, 0, 0], [0, c, -s_], [0, s_, c]], np.float32)
                normals[1] = (normals[1] @ Rx.T).astype(np.float32)
            ids = []
            for k, n in enumerate(tiny["names"]):
                ctx.add_image(900 + k, grays[k], P4s[k])
                ctx.set_depth_map(900 + k, depths[k], kind=1)
                ctx.set_normal_map(900 + k, normals[k], kind=1)
                ids.append(900 + k)
            g = ctx.fuse_points(ids, np.stack(rots), depth_kind=1, inv_proj_mats=np.stack(invs), **kw)
            o = f3.fuse_points(depths, normals, grays, P4s, invs, rots, **kw)
            print("fused points", len(g[0]), kw)
            assert len(g[0]) == len(o[0]) > 0
            assert np.array_equal(g[0], o[0])
            assert np.array_equal(g[1], o[1])
            assert np.array_equal(g[2], o[2])
            small = ctx.fuse_points(ids, np.stack(rots), depth_kind=1, inv_proj_mats=np.stack(invs),
                                    capacity=7, **kw)          # grows to the full count
            assert np.array_equal(small[0], g[0])
            for i in ids:
                ctx.remove_image(i)
    finally:
        ctx.close()


def test_aggregate_3d_dropin(tmp_path_factory):
    """MVS drop-ins, then aggregate_3d.run_fuse on a UTM-anchored site:
    fused.ply (ENU) and aggregate_3d.ply (UTM) written, the DSM GeoTIFF in
    zone 17N close to the UTM ground truth."""
    from vissatsatellitestereo_amd import aggregate_3d, runtime
    from vissatsatellitestereo_amd import colmap_mvs_commands as cmc
    from vissatsatellitestereo_amd.lib.dsm_util import read_dsm_tif
    from vissatsatellitestereo_amd.lib.ply_np_converter import ply2np
    ws = sites.build("geo3d", sites.tiny_spec(), tmp_path_factory.mktemp("g3d"), geo=(30.3, -81.2))
    runtime.reset_context()
    cmc.run_photometric_mvs(ws["mvs"], 4, num_iterations=3)
    cmc.run_consistency_check(ws["mvs"], 4, num_iterations=3)
    pts = aggregate_3d.run_fuse(ws["wd"])
    enu, col, _ = ply2np(os.path.join(ws["mvs"], "fused.ply"))
    utm, col2, comments = ply2np(os.path.join(ws["wd"], "mvs_results", "aggregate_3d", "aggregate_3d.ply"))
    assert len(enu) >= len(utm) == len(pts) > 500
    assert comments == ["projection: UTM 17N"]
    assert np.array_equal(utm, pts)
    img, meta = read_dsm_tif(os.path.join(ws["wd"], "mvs_results", "aggregate_3d", "aggregate_3d_dsm.tif"))
    gt = np.load(os.path.join(ws["wd"], "synthetic_gt", "gt_dsm_utm.npy"))
    ok = ~np.isnan(img) & ~np.isnan(gt)
    err = float(np.median(np.abs(img[ok] - gt[ok])))
    print("aggregate_3d DSM median |dh|", err, "coverage", ok.mean())
    assert (meta["zone_number"], meta["hemisphere"]) == (17, "N")
    assert err < 0.5 and ok.mean() > 0.3
