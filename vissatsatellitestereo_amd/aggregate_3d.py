"""Drop-in for the reference's aggregate_3d.py.

``run_fuse(work_dir)`` (aggregate_3d.py:55-82): 3D fusion of the geometric
depth + normal maps of the dense workspace into a point cloud (the
reference's `colmap stereo_fusion` call, aggregate_3d.py:44-52, with its
min_num_pixels 4, max_reproj_error 2, max_depth_error 1.0,
max_normal_error 10), written to ``colmap/mvs/fused.ply`` (local ENU
x y z, nx ny nz, red green blue); then, as at aggregate_3d.py:64-82, the
points go to UTM (local_to_global + latlon_to_eastnorh, on the device for a
UTM aoi.json; kept in ENU for a local-ENU one), are written to
``mvs_results/aggregate_3d/aggregate_3d.ply`` (double x y z + colour, with
the projection comment) and rasterised into
``mvs_results/aggregate_3d/aggregate_3d_dsm.tif`` by produce_dsm_from_points.

The fusion itself (vsm_fuse_points) is an order-independent restatement of
the stereo_fusion thresholds: a pixel joins the cluster of every image whose
depth map it agrees with in depth (metres), re-projection (pixels) and
normal (degrees), and the first image of a cluster emits its mean; the
depth maps stay on the device when the MVS stages ran in this process.
Keyword extensions: ``depth_type`` ('geometric' as in the reference) and the
four thresholds.
"""
import json
import os

import numpy as np

from . import geodesy, runtime
from .lib.ply_np_converter import np2ply
from .produce_dsm import produce_dsm_from_points


def camera_to_world(P4):
    """R^T of the camera rotation in P4[:3, :3] = K R (K upper triangular with
    a positive diagonal): takes camera-frame normals to the world frame."""
    A = np.asarray(P4, dtype=np.float64)[:3, :3]
    r3 = A[2] / np.linalg.norm(A[2])
    a2 = A[1] - (A[1] @ r3) * r3
    r2 = a2 / np.linalg.norm(a2)
    a1 = A[0] - (A[0] @ r3) * r3 - (A[0] @ r2) * r2
    r1 = a1 / np.linalg.norm(a1)
    return np.stack([r1, r2, r3]).T


def fuse(work_dir, depth_type="geometric", min_num_pixels=4, max_reproj_error=2.0,
         max_depth_error=1.0, max_normal_error=10.0):
    """Device fusion; returns (names, xyz ENU N x 3, normals N x 3, grey N)."""
    mvs = os.path.join(work_dir, "colmap", "mvs")
    ws = runtime.Workspace.get(mvs)
    ctx = runtime.context()
    kind = 1 if depth_type == "geometric" else 0
    ddir = os.path.join(mvs, "stereo", "depth_maps")
    names = [n for n in ws.names if os.path.exists(os.path.join(ddir, f"{n}.{depth_type}.bin"))]
    if not names:
        raise FileNotFoundError(f"no *.{depth_type}.bin depth maps under {ddir}")
    ws.upload(ctx, names)
    for n in names:
        ctx.set_depth_map(ws.ids[n], ws.read_depth(n, depth_type), kind=kind)
        ctx.set_normal_map(ws.ids[n], ws.read_normal(n, depth_type), kind=kind)
    inv = ws.read_inv_proj_mats()
    xyz, nrm, grey = ctx.fuse_points(
        [ws.ids[n] for n in names], np.stack([camera_to_world(ws.P4[n]) for n in names]),
        depth_kind=kind, inv_proj_mats=np.stack([inv[n] for n in names]),
        min_num_pixels=int(min_num_pixels), max_reproj_error=float(max_reproj_error),
        max_depth_error=float(max_depth_error), max_normal_error_deg=float(max_normal_error))
    return names, xyz, nrm, grey


def run_fuse(work_dir, depth_type="geometric", **thresholds):
    names, xyz, nrm, grey = fuse(work_dir, depth_type, **thresholds)
    mvs = os.path.join(work_dir, "colmap", "mvs")
    color = np.repeat(grey[:, None], 3, axis=1)
    np2ply(xyz, os.path.join(mvs, "fused.ply"), color=color, normal=nrm)
    out_dir = os.path.join(work_dir, "mvs_results", "aggregate_3d")
    os.makedirs(out_dir, exist_ok=True)
    aoi = runtime.read_aoi(work_dir)
    geo = geodesy.geo_params_from_aoi(aoi)
    if geo is not None:
        ctx = runtime.context()
        ctx.set_geo_frame(geo)
        try:
            pts = ctx.enu_to_utm(xyz)
        finally:
            ctx.set_geo_frame(None)
        ok = ~np.isnan(pts[:, 0])
        pts, color = pts[ok], color[ok]
        comments = ["projection: UTM {}{}".format(aoi["zone_number"], aoi["hemisphere"])]
    else:
        pts = xyz
        comments = ["projection: local ENU"]
    np2ply(pts, os.path.join(out_dir, "aggregate_3d.ply"), color=color, comments=comments, use_double=True)
    produce_dsm_from_points(work_dir, pts, os.path.join(out_dir, "aggregate_3d_dsm.tif"))
    return pts
