"""Plane-plus-parallax depth reparametrisation (host, FP64).

Drop-in for the reference's ``reparam_depth.py``: same function names,
arguments, output files and numbers.

* ``robust_depth_range`` follows ``reparam_depth.py:40-65``: sort, take the
  ``int(0.02 n)``-th and ``int(0.98 n)``-th values, stretch by -10 / +30,
  fall back to the unstretched pair when the stretched one is empty, and use
  ``(-1e20, -1e20)`` for images without observations.
* ``reparam_depth`` follows ``reparam_depth.py:70-189``: camera depths
  ``(R X + t)_z`` of every track observation with positive depth, the
  reference plane ``z0 = percentile_1(z) - 20``, the per-image 4x4
  ``P4 = [K[R|t]; depth_min * (0, 0, 1, -z0)]``, reparametrised depth
  ``(P4 X)_3 / (P4 X)_2``, and the five text files.  The common depth range
  uses, per point, the depth seen from the *last* image of its track
  (``reparam_depth.py:153-154``) stretched by -10 / +100
  (``reparam_depth.py:179-184``).

The per-observation loops are vectorised per image; results are the same
values the reference's loops produce (FP64 throughout).
"""
import logging
import os

import numpy as np

from .colmap.read_model import read_model, qvec2rotmat, intrinsics


def robust_depth_range(depth_range):
    for img_name in depth_range:
        vals = depth_range[img_name]
        if len(vals) > 0:
            tmp = sorted(vals)
            cnt = len(tmp)
            logging.info("img_name: %s, depth min: %s, max: %s", img_name, tmp[0], tmp[-1])
            min_depth = tmp[int(0.02 * cnt)]
            max_depth = tmp[int(0.98 * cnt)]
            min_new = min_depth - 10
            max_new = max_depth + 30
            if max_new <= min_new:
                min_new, max_new = min_depth, max_depth
            depth_range[img_name] = (min_new, max_new)
        else:
            depth_range[img_name] = (-1e20, -1e20)
    return depth_range


def _observations(images, points3D):
    """Flattened track observations in (point order, track order)."""
    img_ids = sorted(images)
    pos = {iid: k for k, iid in enumerate(img_ids)}
    pt_ids = list(points3D)
    xyz = np.array([points3D[p].xyz for p in pt_ids], dtype=np.float64).reshape(-1, 3)
    obs_pt, obs_img, last = [], [], []
    for k, p in enumerate(pt_ids):
        ids = points3D[p].image_ids
        obs_pt.extend([k] * len(ids))
        obs_img.extend(pos[int(i)] for i in ids)
        last.append(pos[int(ids[-1])] if len(ids) else -1)
    return img_ids, xyz, np.array(obs_pt, dtype=np.int64), np.array(obs_img, dtype=np.int64), \
        np.array(last, dtype=np.int64)


def camera_matrices(cameras, images, image_id):
    im = images[image_id]
    R = qvec2rotmat(im.qvec)
    t = np.asarray(im.tvec, dtype=np.float64).reshape(3, 1)
    K = intrinsics(cameras[im.camera_id])
    return K, R, t


def reparam_depth(sparse_dir, save_dir, camera_model="perspective"):
    assert camera_model in ("perspective", "pinhole")
    cameras, images, points3D = read_model(sparse_dir, ext=".txt")
    img_ids, xyz, obs_pt, obs_img, last = _observations(images, points3D)
    names = [images[i].name for i in img_ids]
    mats = [camera_matrices(cameras, images, i) for i in img_ids]

    # camera depths of all observations (reparam_depth.py:81-93)
    depth_range = {n: [] for n in names}
    for k, (K, R, t) in enumerate(mats):
        sel = obs_pt[obs_img == k]
        if sel.size:
            z = (xyz[sel] @ R[2] + t[2, 0])
            depth_range[names[k]] = list(z[z > 0])
    depth_range = robust_depth_range(depth_range)

    margin = 20.0
    min_z_value = np.percentile(xyz[:, 2], 1) - margin
    logging.info("min_z_value: %s", min_z_value)
    last_row = np.array([0.0, 0.0, 1.0, -min_z_value]).reshape((1, 4))

    P4s = []
    for k, (K, R, t) in enumerate(mats):
        P = K @ np.hstack((R, t))
        P4s.append(np.vstack((P, depth_range[names[k]][0] * last_row)))

    def reparam(k, sel):
        X1 = np.hstack([xyz[sel], np.ones((sel.size, 1))])
        tmp = X1 @ P4s[k].T
        return tmp[:, 3] / tmp[:, 2]

    reparam_range = {n: [] for n in names}
    last_rows = {}
    for k in range(len(names)):
        sel = obs_pt[obs_img == k]
        if sel.size:
            last_rows[names[k]] = depth_range[names[k]][0] * last_row
            d = reparam(k, sel)
            reparam_range[names[k]] = list(d[d > 0])
    common = []
    has = last >= 0
    for k in range(len(names)):
        sel = np.nonzero(has & (last == k))[0]
        if sel.size:
            d = reparam(k, sel)
            common.extend(d[d > 0])
    reparam_range = robust_depth_range(reparam_range)

    with open(os.path.join(save_dir, "raw_depth.txt"), "w") as fp:
        fp.write("# format: img_name, depth_min, depth_max\n")
        for n in sorted(depth_range):
            fp.write("{} {} {}\n".format(n, *map(float, depth_range[n])))
    with open(os.path.join(save_dir, "reparam_depth.txt"), "w") as fp:
        fp.write("# format: img_name, depth_min, depth_max\n")
        for n in sorted(reparam_range):
            fp.write("{} {} {}\n".format(n, *map(float, reparam_range[n])))
    with open(os.path.join(save_dir, "last_rows.txt"), "w") as fp:
        for n in sorted(last_rows):
            v = last_rows[n]
            fp.write("{} {} {} {} {}\n".format(n, float(v[0, 0]), float(v[0, 1]),
                                               float(v[0, 2]), float(v[0, 3])))
    with open(os.path.join(save_dir, "reference_plane.txt"), "w") as fp:
        fp.write("{} {} {} {}\n".format(*map(float, last_row[0])))

    common = sorted(common)
    cnt = len(common)
    min_depth = common[int(0.02 * cnt)] - 10
    max_depth = common[int(0.98 * cnt)] + 100.0
    logging.info("%d points, depth_min: %s, depth_max: %s", cnt, min_depth, max_depth)
    with open(os.path.join(save_dir, "depth_ranges.txt"), "w") as fp:
        for n in sorted(last_rows):
            fp.write("{} {} {}\n".format(n, float(min_depth), float(max_depth)))


# ----------------------------------------------------------------------------
# Readers for the files above (used by the MVS and fusion drop-ins).
# ----------------------------------------------------------------------------
def read_name_values(path, skip_comments=True):
    out = {}
    with open(path) as fp:
        for line in fp:
            s = line.strip()
            if not s or (skip_comments and s[0] == "#"):
                continue
            e = s.split()
            out[e[0]] = [float(v) for v in e[1:]]
    return out


def projection_matrices(mvs_dir):
    """Per-image FP64 4x4 ``P4 = [K[R|t]; last_row]`` of the dense workspace
    (``reparam_depth.py:136-141`` with ``last_rows.txt``)."""
    cameras, images, _ = read_model(os.path.join(mvs_dir, "sparse"), ext=".txt")
    last_rows = read_name_values(os.path.join(mvs_dir, "last_rows.txt"))
    out = {}
    for iid, im in images.items():
        if im.name not in last_rows:
            continue
        K, R, t = camera_matrices(cameras, images, iid)
        P = K @ np.hstack((R, t))
        out[im.name] = np.vstack((P, np.array(last_rows[im.name]).reshape(1, 4)))
    return out
