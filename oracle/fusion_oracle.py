"""CPU restatement of the reference's 2.5D fusion path (numpy).

TEST INFRASTRUCTURE (the oracle): imported only by tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg, as the checker.

Follows, function by function:
* ``read_array``              colmap/read_dense.py:36-51
* ``backproject``             aggregate_2p5d_util.py:75-93 (d <= 0 -> NaN,
                              X = inv(P4) [col, row, 1, d], X[:3] /= X[3]);
                              the 4x4 product is written out term by term
                              ((m0 col + m1 row) + m2) + m3 d so the FP64
                              operation order is fixed (np.dot's BLAS order is
                              not, aggregate_2p5d_util.py:87)
* ``proj_to_grid``            lib/proj_to_grid.py:41-81 (floor cell index,
                              bounds mask, per-cell nanmax, NaN cells filled
                              with np.median of the in-bounds non-NaN 3x3
                              neighbours of the unfilled raster)
* ``median_blur3``            the 3x3 median of produce_dsm.py:58 and
                              aggregate_2p5d.py:81, with replicated border;
                              NaN handling of cv2.medianBlur is unpinned (cv2
                              is not installed), so this build's rule is: NaN
                              excluded, even count -> float32 mean of the two
                              middle values, all-NaN -> NaN
* ``produce_dsm_from_points`` produce_dsm.py:46-58
* ``fuse``                    aggregate_2p5d.py:66-81 (<= min_measurements-1
                              valid -> NaN, nanmedian, MAD, |x-med| > MAD
                              dropped, mean, 3x3 median); the mean accumulates
                              in FP64 (the reference's float32 nanmean order
                              is numpy-internal)
* ``grid_from_aoi``           produce_dsm.py:51-55
* ``view_dsm(..., frame)``    with a geodetic frame (30 constants of
                              vsm_geo_frame) the back-projected ENU points go
                              through local_to_global + latlon_to_eastnorh
                              (aggregate_2p5d_util.py:95-98, geo_oracle.c)
                              before the grid, which is then aoi.json's UTM
                              grid; points outside the series' domain drop out
"""
import json
import math
import os

import numpy as np


def read_array(path):
    with open(path, "rb") as fid:
        raw = fid.read()
    amp = 0
    pos = 0
    while amp < 3:
        if raw[pos:pos + 1] == b"&":
            amp += 1
        pos += 1
    w, h, c = (int(v) for v in raw[:pos].decode().split("&")[:3])
    arr = np.frombuffer(raw[pos:], dtype="<f4")
    arr = arr.reshape((w, h, c), order="F")
    return np.transpose(arr, (1, 0, 2)).squeeze()


def backproject(depth, inv_proj):
    """ENU points (N x 3, FP64) of the valid pixels, in row-major pixel order."""
    depth = np.asarray(depth, dtype=np.float32)
    h, w = depth.shape
    d = depth.astype(np.float64).reshape(-1)
    d[~(d > 0)] = np.nan
    row, col = np.meshgrid(np.arange(h, dtype=np.float64), np.arange(w, dtype=np.float64), indexing="ij")
    row = row.reshape(-1)
    col = col.reshape(-1)
    M = np.asarray(inv_proj, dtype=np.float64).reshape(4, 4)
    X = [((M[i, 0] * col + M[i, 1] * row) + M[i, 2]) + M[i, 3] * d for i in range(4)]
    E, N, U = X[0] / X[3], X[1] / X[3], X[2] / X[3]
    ok = np.isfinite(E) & np.isfinite(N) & np.isfinite(U)
    return np.stack([E[ok], N[ok], U[ok]], axis=1)


def _median_of_sorted(s, n, half_dtype):
    """Median of the first n entries of each sorted row (n >= 1)."""
    lo = np.take_along_axis(s, ((n - 1) // 2)[..., None], axis=-1)[..., 0]
    hi = np.take_along_axis(s, (n // 2)[..., None], axis=-1)[..., 0]
    even = (lo + hi) * half_dtype(0.5)
    return np.where(n % 2 == 1, hi, even)


def proj_to_grid(points, xoff, yoff, xresolution, yresolution, xsize, ysize, propagate=False):
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    row = np.floor((yoff - pts[:, 1]) / xresolution)
    col = np.floor((pts[:, 0] - xoff) / yresolution)
    val = pts[:, 2]
    m = (row >= 0) & (col >= 0) & (row < ysize) & (col < xsize) & ~np.isnan(val)
    idx = (row[m] * xsize + col[m]).astype(np.int64)
    flat = np.full(xsize * ysize, -np.inf)
    np.maximum.at(flat, idx, val[m])
    flat[flat == -np.inf] = np.nan
    dsm = flat.reshape(ysize, xsize)
    # hole fill from the unfilled raster, in-bounds neighbours only
    pad = np.pad(dsm, 1, mode="constant", constant_values=np.nan)
    nb = np.stack([pad[1 + dr:1 + dr + ysize, 1 + dc:1 + dc + xsize]
                   for dr in (-1, 0, 1) for dc in (-1, 0, 1)], axis=-1)
    n = np.sum(~np.isnan(nb), axis=-1)
    s = np.sort(nb, axis=-1)
    med = _median_of_sorted(s, np.maximum(n, 1), np.float64)
    out = dsm.copy()
    fill = np.isnan(dsm) & (n > 0)
    out[fill] = med[fill]
    return out


def median_blur3(img):
    a = np.asarray(img, dtype=np.float32)
    h, w = a.shape
    pad = np.pad(a, 1, mode="edge")
    nb = np.stack([pad[1 + dr:1 + dr + h, 1 + dc:1 + dc + w]
                   for dr in (-1, 0, 1) for dc in (-1, 0, 1)], axis=-1)
    n = np.sum(~np.isnan(nb), axis=-1)
    s = np.sort(nb, axis=-1)
    med = _median_of_sorted(s, np.maximum(n, 1), np.float32).astype(np.float32)
    med[n == 0] = np.nan
    return med


def grid_from_aoi(aoi, e_resolution=0.5, n_resolution=0.5):
    e_size = int(aoi["width"] / e_resolution) + 1
    n_size = int(aoi["height"] / n_resolution) + 1
    return (aoi["ul_easting"], aoi["ul_northing"], e_resolution, n_resolution, e_size, n_size)


def produce_dsm_from_points(points, grid):
    ul_e, ul_n, e_res, n_res, e_size, n_size = grid
    dsm = proj_to_grid(points, ul_e, ul_n, e_res, n_res, e_size, n_size, propagate=True)
    return median_blur3(dsm.astype(np.float32))


def view_dsm(depth, inv_proj, grid, frame=None):
    pts = backproject(depth, inv_proj)
    if frame is not None:
        import geo_oracle
        pts = geo_oracle.enu_to_utm(frame, pts)
        pts = pts[~np.isnan(pts[:, 0])]
    return produce_dsm_from_points(pts, grid)


def fuse(stack, min_measurements=3):
    """stack: V x H x W float32 per-view DSMs (NaN = empty)."""
    a = np.asarray(stack, dtype=np.float32)
    a = np.moveaxis(a, 0, -1).copy()           # H x W x V
    n = np.sum(~np.isnan(a), axis=-1)
    a[n < min_measurements] = np.nan
    n = np.where(n < min_measurements, 0, n)
    s = np.sort(a, axis=-1)
    med = _median_of_sorted(s, np.maximum(n, 1), np.float32).astype(np.float32)
    dev = np.abs(a - med[..., None])
    sd = np.sort(dev, axis=-1)
    mad = _median_of_sorted(sd, np.maximum(n, 1), np.float32).astype(np.float32)
    keep = ~np.isnan(a) & ~(dev > mad[..., None])
    cnt = keep.sum(-1)
    tot = np.zeros(a.shape[:2])
    for k in range(a.shape[-1]):          # sequential FP64 accumulation in view order
        tot = tot + np.where(keep[..., k], a[..., k].astype(np.float64), 0.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        mean = np.where(cnt > 0, tot / np.maximum(cnt, 1), np.nan).astype(np.float32)
    mean[n == 0] = np.nan
    return median_blur3(mean)


def read_inv_proj_mats(path):
    """aggregate_2p5d_util.py:55-61."""
    out = {}
    with open(path) as fp:
        for line in fp.readlines():
            tmp = line.split(" ")
            out[tmp[0]] = np.array([float(tmp[i]) for i in range(1, 17)]).reshape((4, 4))
    return out


def geo_frame_from_aoi(aoi):
    """Frame constants of a UTM aoi.json (None for a local-ENU one)."""
    import geo_oracle
    if int(aoi.get("zone_number", 0)) == 0 or str(aoi.get("frame", "")).startswith("local"):
        return None
    return geo_oracle.frame_constants((aoi["lat_min"] + aoi["lat_max"]) / 2.0,
                                      (aoi["lon_min"] + aoi["lon_max"]) / 2.0, float(aoi["alt_min"]),
                                      int(aoi["zone_number"]), 0 if aoi["hemisphere"] == "N" else 1)


def run_fuse(work_dir, depth_type="geometric", min_measurements=3, resolution=0.5, frame="aoi"):
    """Whole-site fusion from the workspace files; returns (fused, per_view).
    frame: "aoi" derives the geodetic frame from aoi.json, else the 30
    constants or None (local ENU grid)."""
    mvs = os.path.join(work_dir, "colmap", "mvs")
    inv = read_inv_proj_mats(os.path.join(mvs, "inv_proj_mats.txt"))
    with open(os.path.join(work_dir, "aoi.json")) as fp:
        aoi = json.load(fp)
    grid = grid_from_aoi(aoi, resolution, resolution)
    if isinstance(frame, str):
        frame = geo_frame_from_aoi(aoi)
    ddir = os.path.join(mvs, "stereo", "depth_maps")
    items = sorted(i for i in os.listdir(ddir) if depth_type in i)
    per_view = []
    for item in items:
        name = item[:item.rfind(f".{depth_type}.bin")]
        per_view.append(view_dsm(read_array(os.path.join(ddir, item)), inv[name], grid, frame))
    stack = np.stack(per_view)
    return fuse(stack, min_measurements), stack


def median_abs_error(dsm, gt):
    ok = ~np.isnan(dsm) & ~np.isnan(gt)
    if not ok.any():
        return math.nan, 0.0
    return float(np.median(np.abs(dsm[ok] - gt[ok]))), float(ok.mean())
