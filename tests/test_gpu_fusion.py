"""Device 2.5D fusion against the fusion oracle (oracle/fusion_oracle.py).

Integer grid indexing and every raster value must be bit-exact: the device
back-projects in FP64 with the oracle's operation order, takes the per-cell
max of FP64 heights, fills holes with the FP64 median and runs the float32
3x3 median with the same rule; the cross-view fusion accumulates its mean in
FP64 in view order as the oracle does.
"""
import json
import os

import numpy as np
import pytest

import sites
import fusion_oracle as fo

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "fusion_golden.json")


@pytest.fixture(scope="module")
def ctx():
    from vissatsatellitestereo_amd.mvs import MvsContext
    c = MvsContext(device=0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def tiny(site_dir):
    return sites.build("tiny", sites.tiny_spec(), site_dir)


def _same(a, b):
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


def test_proj_to_grid_bit_exact(ctx):
    rng = np.random.default_rng(3)
    for n, size in [(0, (5, 4)), (7, (6, 6)), (5000, (41, 37)), (200000, (301, 257))]:
        pts = np.column_stack([rng.uniform(-3, size[1] * 0.5 + 3, n), rng.uniform(-size[0] * 0.5 - 3, 3, n),
                               rng.normal(20, 5, n)])
        if n > 10:
            pts[::97, 2] = np.nan
        grid = (0.0, 0.0, 0.5, 0.5, size[1], size[0])
        g = ctx.proj_to_grid(pts, grid)
        o = fo.proj_to_grid(pts, *grid)
        assert _same(g, o), n


def test_golden_cases_on_device(ctx):
    gold = json.load(open(GOLD))
    for case in gold["proj_to_grid"]:
        g = ctx.proj_to_grid(np.array(case["points"], dtype=np.float64).reshape(-1, 3), tuple(case["grid"]))
        assert _same(g, np.array(case["expected"], dtype=np.float64)), case["name"]
    for case in gold["median_blur3"]:
        g = ctx.median_blur3(np.array(case["input"], dtype=np.float32))
        assert _same(g, np.array(case["expected"], dtype=np.float32)), case["name"]


def test_median_blur_bit_exact(ctx):
    rng = np.random.default_rng(5)
    for shape in [(1, 1), (1, 9), (7, 1), (64, 80), (513, 257)]:
        a = rng.normal(0, 10, shape).astype(np.float32)
        a[rng.random(shape) < 0.3] = np.nan
        assert _same(ctx.median_blur3(a), fo.median_blur3(a)), shape


def test_fuse_dsm_bit_exact(ctx, tiny):
    """Per-view and fused DSMs of noisy ground-truth depth maps (with holes)."""
    names = tiny["names"]
    rng = np.random.default_rng(11)
    grid = fo.grid_from_aoi(json.load(open(os.path.join(tiny["wd"], "aoi.json"))))
    ids, invs, depths = [], [], []
    for i, n in enumerate(names):
        d = sites.gt_depth(tiny, n).astype(np.float32)
        d = d + rng.normal(0, 0.3, d.shape).astype(np.float32)
        d[rng.random(d.shape) < 0.1] = 0.0
        d[:5, :] = -1.0
        h, w = d.shape
        ctx.register_geometry(100 + i, w, h, tiny["P4"][n])
        ctx.set_depth_map(100 + i, d, kind=1)
        ids.append(100 + i)
        invs.append(np.linalg.inv(tiny["P4"][n]))
        depths.append(d)
    for mm in (3, 2):
        fused, pv = ctx.fuse_dsm(grid, ids, depth_kind=1, min_measurements=mm, inv_proj_mats=np.stack(invs),
                                 per_view=True)
        views = [fo.view_dsm(d, inv, grid) for d, inv in zip(depths, invs)]
        for k in range(len(ids)):
            assert _same(pv[k], views[k]), (mm, k)
        ofused = fo.fuse(np.stack(views), mm)
        assert _same(fused, ofused), mm
    gt = np.load(os.path.join(tiny["wd"], "synthetic_gt", "gt_dsm.npy"))
    err, cov = fo.median_abs_error(fused, gt)
    print("fused DSM of noisy GT depths: median |dh|", err, "coverage", cov)
    assert err < 0.3 and cov > 0.5


@pytest.mark.parametrize("shape", [(1, 1), (1, 7), (9, 1), (17, 65), (33, 130)])
def test_fuse_dsm_odd_grids_bit_exact(ctx, tiny, shape):
    """Per-view DSMs on grids that are one cell thin or not a multiple of the
    per-view kernel's 64 x 16-cell tile (single rows and columns take the
    replicated border on both sides), against the oracle; the grids sit over
    the middle of the site at 0.37 m cells, so cells are filled, empty and
    hole-filled."""
    n_size, e_size = shape
    aoi = json.load(open(os.path.join(tiny["wd"], "aoi.json")))
    ul_e = aoi["ul_easting"] + 0.4 * aoi["width"]
    ul_n = aoi["ul_northing"] - 0.4 * aoi["height"]
    grid = (ul_e, ul_n, 0.37, 0.37, e_size, n_size)
    names = tiny["names"]
    rng = np.random.default_rng(5)
    ids, invs, depths = [], [], []
    for i, n in enumerate(names):
        d = sites.gt_depth(tiny, n).astype(np.float32)
        d[rng.random(d.shape) < 0.3] = 0.0
        h, w = d.shape
        ctx.register_geometry(200 + i, w, h, tiny["P4"][n])
        ctx.set_depth_map(200 + i, d, kind=1)
        ids.append(200 + i)
        invs.append(np.linalg.inv(tiny["P4"][n]))
        depths.append(d)
    fused, pv = ctx.fuse_dsm(grid, ids, depth_kind=1, min_measurements=2, inv_proj_mats=np.stack(invs),
                             per_view=True)
    views = [fo.view_dsm(d, inv, grid) for d, inv in zip(depths, invs)]
    assert fused.shape == shape
    for k in range(len(ids)):
        assert _same(pv[k], views[k]), k
    assert _same(fused, fo.fuse(np.stack(views), 2))
    if n_size * e_size > 1:                     # the single cell lies in a gap of every view
        assert any(np.isfinite(v).any() for v in views)


def test_fuse_dsm_rejects_bad_arguments(ctx):
    from vissatsatellitestereo_amd._lib import VsmError
    with pytest.raises(VsmError):
        ctx.fuse_dsm((0.0, 0.0, 0.5, 0.5, 0, 10), [100])
    with pytest.raises(VsmError):
        ctx.fuse_dsm((0.0, 0.0, 0.5, 0.5, 10, 10), [987654])
