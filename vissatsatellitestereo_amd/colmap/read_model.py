"""COLMAP sparse text model reader.

Same types and field meaning as the reference's ``colmap/read_model.py``:
``Camera`` (``read_model.py:41-42``), ``Image`` (``read_model.py:43-44``),
``Point3D`` (``read_model.py:45-46``); text readers follow
``read_cameras_text`` (``read_model.py:83-106``), ``read_images_text``
(``read_model.py:138-166``) and ``read_points3D_text``
(``read_model.py:204-228``); ``read_model(path, ext)`` as
``read_model.py:261-270`` (text only: the MVS stage reads ``.txt``,
``reparam_depth.py:73``).
"""
import collections
import os

import numpy as np

Camera = collections.namedtuple("Camera", ["id", "model", "width", "height", "params"])
Image = collections.namedtuple("Image", ["id", "qvec", "tvec", "camera_id", "name",
                                         "xys", "point3D_ids"])
Point3D = collections.namedtuple("Point3D", ["id", "xyz", "rgb", "error", "image_ids",
                                             "point2D_idxs"])


def _data_lines(path):
    with open(path, "r") as fid:
        for line in fid:
            s = line.strip()
            if s and s[0] != "#":
                yield s


def read_cameras_text(path):
    cameras = {}
    for s in _data_lines(path):
        e = s.split()
        cid = int(e[0])
        cameras[cid] = Camera(id=cid, model=e[1], width=int(e[2]), height=int(e[3]),
                              params=np.array([float(v) for v in e[4:]]))
    return cameras


def read_images_text(path):
    """Two lines per image; the second (2D points) may be empty."""
    images = {}
    with open(path, "r") as fid:
        while True:
            line = fid.readline()
            if not line:
                break
            s = line.strip()
            if not s or s[0] == "#":
                continue
            e = s.split()
            iid = int(e[0])
            qvec = np.array([float(v) for v in e[1:5]])
            tvec = np.array([float(v) for v in e[5:8]])
            cam_id = int(e[8])
            name = e[9]
            pts = fid.readline().split()
            xys = np.column_stack([np.array([float(v) for v in pts[0::3]]),
                                   np.array([float(v) for v in pts[1::3]])]) if pts else np.zeros((0, 2))
            p3d = np.array([int(v) for v in pts[2::3]], dtype=np.int64)
            images[iid] = Image(id=iid, qvec=qvec, tvec=tvec, camera_id=cam_id, name=name,
                                xys=xys, point3D_ids=p3d)
    return images


def read_points3D_text(path):
    points = {}
    for s in _data_lines(path):
        e = s.split()
        pid = int(e[0])
        points[pid] = Point3D(id=pid, xyz=np.array([float(v) for v in e[1:4]]),
                              rgb=np.array([int(v) for v in e[4:7]]), error=float(e[7]),
                              image_ids=np.array([int(v) for v in e[8::2]], dtype=np.int64),
                              point2D_idxs=np.array([int(v) for v in e[9::2]], dtype=np.int64))
    return points


def read_model(path, ext=".txt"):
    if ext != ".txt":
        raise ValueError("only the text model is supported (the MVS stage reads .txt)")
    cameras = read_cameras_text(os.path.join(path, "cameras" + ext))
    images = read_images_text(os.path.join(path, "images" + ext))
    points3D = read_points3D_text(os.path.join(path, "points3D" + ext))
    return cameras, images, points3D


def qvec2rotmat(qvec):
    """Rotation of a (w, x, y, z) quaternion after normalisation (the
    pyquaternion ``rotation_matrix`` used at ``reparam_depth.py:89``
    normalises first)."""
    w, x, y, z = np.asarray(qvec, dtype=np.float64) / np.linalg.norm(qvec)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def intrinsics(camera):
    """K of a PERSPECTIVE (fx, fy, cx, cy, s) or PINHOLE camera
    (``reparam_depth.py:125-134``)."""
    p = camera.params
    if camera.model == "PINHOLE" or len(p) == 4:
        fx, fy, cx, cy = p[:4]
        s = 0.0
    else:
        fx, fy, cx, cy, s = p[:5]
    return np.array([[fx, s, cx], [0.0, fy, cy], [0.0, 0.0, 1.0]])
