"""Line-by-line CPU restatement of the reference's reparam_depth.py.

TEST INFRASTRUCTURE (the oracle): imported only by tests/ as the checker of
the vectorised product implementation (vissatsatellitestereo_amd/reparam_depth.py).

Pure-Python loops over points x track views as at reparam_depth.py:81-93 and
114-154; robust_depth_range as reparam_depth.py:40-65; output text as
reparam_depth.py:159-189.  Quaternions are turned into rotations with the
normalised unit-quaternion formula (the pyquaternion rotation_matrix used at
reparam_depth.py:89 normalises first).
"""
import math
import os


def _rot(q):
    w, x, y, z = q
    n = math.sqrt(w * w + x * x + y * y + z * z)
    w, x, y, z = w / n, x / n, y / n, z / n
    return [[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
            [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
            [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]]


def _read(sparse_dir):
    cams, imgs, pts = {}, {}, {}
    for line in open(os.path.join(sparse_dir, "cameras.txt")):
        s = line.strip()
        if s and s[0] != "#":
            e = s.split()
            cams[int(e[0])] = [float(v) for v in e[4:]]
    lines = [ln for ln in open(os.path.join(sparse_dir, "images.txt"))]
    i = 0
    while i < len(lines):
        s = lines[i].strip()
        if not s or s[0] == "#":
            i += 1
            continue
        e = s.split()
        imgs[int(e[0])] = dict(q=[float(v) for v in e[1:5]], t=[float(v) for v in e[5:8]],
                               cam=int(e[8]), name=e[9])
        i += 2
    for line in open(os.path.join(sparse_dir, "points3D.txt")):
        s = line.strip()
        if s and s[0] != "#":
            e = s.split()
            pts[int(e[0])] = dict(xyz=[float(v) for v in e[1:4]], ids=[int(v) for v in e[8::2]])
    return cams, imgs, pts


def robust_depth_range(depth_range):
    for img_name in depth_range:
        if depth_range[img_name]:
            tmp = sorted(depth_range[img_name])
            cnt = len(tmp)
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


def _percentile(vals, q):
    """numpy's default (linear) percentile."""
    s = sorted(vals)
    pos = (len(s) - 1) * q / 100.0
    lo = int(math.floor(pos))
    hi = min(lo + 1, len(s) - 1)
    return s[lo] + (s[hi] - s[lo]) * (pos - lo)


def reparam_depth(sparse_dir):
    """Returns the five tables the reference writes, as Python values."""
    cams, imgs, pts = _read(sparse_dir)
    depth_range = {im["name"]: [] for im in imgs.values()}
    z_values = []
    for p in pts.values():
        x = p["xyz"]
        z_values.append(x[2])
        for iid in p["ids"]:
            im = imgs[iid]
            R = _rot(im["q"])
            depth = sum(R[2][k] * x[k] for k in range(3)) + im["t"][2]
            if depth > 0:
                depth_range[im["name"]].append(depth)
    depth_range = robust_depth_range(depth_range)
    min_z = _percentile(z_values, 1) - 20.0
    last_row = [0.0, 0.0, 1.0, -min_z]
    last_rows = {}
    reparam = {im["name"]: [] for im in imgs.values()}
    common = []
    for p in pts.values():
        x = p["xyz"] + [1.0]
        depth = 0
        for iid in p["ids"]:
            im = imgs[iid]
            R = _rot(im["q"])
            fx, fy, cx, cy, s = cams[im["cam"]][:5]
            K = [[fx, s, cx], [0.0, fy, cy], [0.0, 0.0, 1.0]]
            Rt = [R[r] + [im["t"][r]] for r in range(3)]
            P = [[sum(K[r][k] * Rt[k][c] for k in range(3)) for c in range(4)] for r in range(3)]
            dmin = depth_range[im["name"]][0]
            P4 = P + [[dmin * v for v in last_row]]
            if im["name"] not in last_rows:
                last_rows[im["name"]] = [dmin * v for v in last_row]
            tmp = [sum(P4[r][c] * x[c] for c in range(4)) for r in range(4)]
            depth = tmp[3] / tmp[2]
            if depth > 0:
                reparam[im["name"]].append(depth)
        if depth > 0:
            common.append(depth)
    reparam = robust_depth_range(reparam)
    common = sorted(common)
    cnt = len(common)
    rng = (common[int(0.02 * cnt)] - 10, common[int(0.98 * cnt)] + 100.0)
    return dict(raw_depth=depth_range, reparam_depth=reparam, last_rows=last_rows,
                reference_plane=last_row, depth_ranges={n: rng for n in last_rows})
