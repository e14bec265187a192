"""Drop-in for the reference's aggregate_2p5d_util.py.

``convert_depth_maps(work_dir, out_dir, depth_type, max_processes=-1)``
(aggregate_2p5d_util.py:125-146) turns every ``*.<depth_type>.bin`` depth map
into a per-view DSM written to ``out_dir/dsm_tif/<name>.tif``: back-projection
through inv(P4) from inv_proj_mats.txt (aggregate_2p5d_util.py:55-93),
per-cell nanmax with hole fill (lib/proj_to_grid.py:41-81) and 3x3 median
(produce_dsm.py:58), all in one device pass per view instead of a process
pool.  For a UTM aoi.json (the reference's, stereo_pipeline.py:185-226) every
back-projected ENU point goes through local_to_global + latlon_to_eastnorh
(aggregate_2p5d_util.py:95-98) on the device before it is rasterised, so the
DSM is aoi.json's UTM grid of WGS84 altitudes; a local-ENU aoi.json
(zone_number 0, the synthetic sites) keeps the grid in the ENU frame.  The
per-view DSMs are written as GeoTIFFs with nodata -10000 (produce_dsm.py:59-61,
lib/dsm_util.py:122-159).  max_processes is accepted for signature
compatibility; the GPU needs no worker pool.
"""
import os
import shutil

import numpy as np

from . import geodesy, runtime


def _items(mvs_dir, depth_type):
    ddir = os.path.join(mvs_dir, "stereo", "depth_maps")
    out = []
    for item in sorted(os.listdir(ddir)):
        if depth_type not in item:
            continue
        idx = item.rfind(".{}.bin".format(depth_type))
        if idx == -1:
            continue
        out.append((item, item[:idx]))
    return out


def per_view_dsms(work_dir, depth_type="geometric", min_measurements=3, resolution=0.5):
    """Device fusion of the workspace's depth maps.

    Returns (names, fused DSM, per-view DSMs V x n_size x e_size)."""
    mvs_dir = os.path.join(work_dir, "colmap", "mvs")
    ws = runtime.Workspace.get(mvs_dir)
    ctx = runtime.context()
    kind = 1 if depth_type == "geometric" else 0
    items = _items(mvs_dir, depth_type)
    if not items:
        raise FileNotFoundError(f"no *.{depth_type}.bin depth maps under {mvs_dir}")
    inv = ws.read_inv_proj_mats()
    names = [n for _, n in items]
    for item, name in items:
        depth = ws.read_depth(name, depth_type)
        ws.register(ctx, name, depth.shape)
        ctx.set_depth_map(ws.ids[name], depth, kind=kind)
    aoi = runtime.read_aoi(work_dir)
    grid = runtime.dsm_grid(aoi, resolution, resolution)
    mats = np.stack([inv[n] for n in names])
    ctx.set_geo_frame(geodesy.geo_params_from_aoi(aoi))
    try:
        fused, pv = ctx.fuse_dsm(grid, [ws.ids[n] for n in names], depth_kind=kind,
                                 min_measurements=min_measurements, inv_proj_mats=mats, per_view=True)
    finally:
        ctx.set_geo_frame(None)
    return names, fused, pv


def convert_depth_maps(work_dir, out_dir, depth_type, max_processes=-1, min_measurements=3):
    if os.path.exists(out_dir):
        shutil.rmtree(out_dir)
    for sub in (out_dir, os.path.join(out_dir, "dsm_tif")):
        os.makedirs(sub, exist_ok=True)
    names, fused, pv = per_view_dsms(work_dir, depth_type, min_measurements)
    aoi = runtime.read_aoi(work_dir)
    for name, dsm in zip(names, pv):
        runtime.write_dsm_tif_aoi(dsm, os.path.join(out_dir, "dsm_tif", name[:-4] + ".tif"), aoi)
    return names, fused, pv
