"""The reference-surface drop-ins on a dense workspace, on the GPU:
run_photometric_mvs -> run_consistency_check -> aggregate_2p5d.run_fuse,
checked against the fusion oracle on the files they wrote (bit-exact) and
against the synthetic ground truth."""
import os

import numpy as np
import pytest

import sites

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ws(tmp_path_factory):
    root = tmp_path_factory.mktemp("dropin")
    return sites.build("tiny", sites.tiny_spec(), root)


def test_pipeline_files_and_fusion(ws):
    from vissatsatellitestereo_amd import colmap_mvs_commands as cmc
    from vissatsatellitestereo_amd import aggregate_2p5d, runtime
    from vissatsatellitestereo_amd.colmap.read_dense import read_array
    import fusion_oracle as fo
    runtime.reset_context()
    mvs = ws["mvs"]
    ph = cmc.run_photometric_mvs(mvs, 4, num_iterations=3)
    for n in ws["names"]:
        d = read_array(os.path.join(mvs, "stereo", "depth_maps", f"{n}.photometric.bin"))
        nm = read_array(os.path.join(mvs, "stereo", "normal_maps", f"{n}.photometric.bin"))
        assert d.shape == ws["imgs"][n].shape and nm.shape == d.shape + (3,)
        assert np.array_equal(d, ph[n]["depth"])
    assert os.path.exists(os.path.join(mvs, "inv_proj_mats.txt"))
    geo = cmc.run_consistency_check(mvs, 4, num_iterations=3)
    for n in ws["names"]:
        d = read_array(os.path.join(mvs, "stereo", "depth_maps", f"{n}.geometric.bin"))
        assert np.array_equal(d, geo[n]["depth"])
        assert os.path.exists(os.path.join(mvs, "stereo", "consistency_graphs", f"{n}.geometric.npy"))
        from vissatsatellitestereo_amd.colmap import consistency_graph
        g = consistency_graph.read(os.path.join(mvs, "stereo", "consistency_graphs", f"{n}.geometric.bin"))
        m = np.load(os.path.join(mvs, "stereo", "consistency_graphs", f"{n}.geometric.npy"))
        assert set(g) == set(zip(*np.nonzero(m)))
        assert all(len(v) == bin(int(m[k])).count("1") for k, v in g.items())
        gt = sites.gt_depth(ws, n)
        valid = d > 0
        assert valid.mean() > 0.3
        assert np.median(np.abs(d - gt)[valid]) < 0.6
    fused = aggregate_2p5d.run_fuse(ws["wd"], min_measurements=3)
    ofused, ostack = fo.run_fuse(ws["wd"], "geometric", 3)
    assert np.array_equal(np.isnan(fused), np.isnan(ofused))
    assert np.array_equal(fused[~np.isnan(fused)], ofused[~np.isnan(ofused)])
    tif = os.path.join(ws["wd"], "mvs_results", "aggregate_2p5d", "aggregate_2p5d_dsm.tif")
    assert np.array_equal(runtime.read_tif(tif), fused, equal_nan=True)
    for n, pv in zip(sorted(ws["names"]), ostack):
        t = runtime.read_tif(os.path.join(ws["wd"], "colmap", "mvs", "dsm", "dsm_tif", n[:-4] + ".tif"))
        assert np.array_equal(t, pv, equal_nan=True)
    gt = np.load(os.path.join(ws["wd"], "synthetic_gt", "gt_dsm.npy"))
    err, cov = fo.median_abs_error(fused, gt)
    print("drop-in fused DSM median |dh|", err, "coverage", cov)
    assert err < 0.5 and cov > 0.3


def test_consistency_check_resumes_from_files(ws):
    """A fresh context (new process in the reference) reads the photometric
    maps back from the workspace (use_exist_photom) and reaches the same
    accuracy."""
    from vissatsatellitestereo_amd import colmap_mvs_commands as cmc
    from vissatsatellitestereo_amd import runtime
    runtime.reset_context()
    mvs = ws["mvs"]
    geo = cmc.run_consistency_check(mvs, 4, num_iterations=3)
    for n in ws["names"]:
        gt = sites.gt_depth(ws, n)
        d = geo[n]["depth"]
        valid = d > 0
        assert valid.mean() > 0.3
        assert np.median(np.abs(d - gt)[valid]) < 0.6


def test_proj_to_grid_and_produce_dsm_dropins(ws, tmp_path):
    from vissatsatellitestereo_amd.lib.proj_to_grid import proj_to_grid
    from vissatsatellitestereo_amd import produce_dsm, runtime
    import fusion_oracle as fo
    rng = np.random.default_rng(2)
    pts = np.column_stack([rng.uniform(-20, 20, 3000), rng.uniform(-20, 20, 3000), rng.normal(5, 2, 3000)])
    a = proj_to_grid(pts, -25.0, 25.0, 0.5, 0.5, 101, 101)
    b = fo.proj_to_grid(pts, -25.0, 25.0, 0.5, 0.5, 101, 101)
    assert np.array_equal(a, b, equal_nan=True)
    out = tmp_path / "d.tif"
    g = produce_dsm.produce_dsm_from_points(ws["wd"], pts, str(out))
    aoi = runtime.read_aoi(ws["wd"])
    assert g[0] == aoi["ul_easting"] and g[2] == int(aoi["width"] / 0.5) + 1
    exp = fo.produce_dsm_from_points(pts, fo.grid_from_aoi(aoi))
    assert np.array_equal(runtime.read_tif(str(out)), exp, equal_nan=True)
