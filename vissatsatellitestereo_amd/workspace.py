"""COLMAP dense workspace: layout, config files and source-view selection.

``prepare_dense_workspace`` is ``StereoPipeline.run_reparam_depth``
(``stereo_pipeline.py:379-442``) without the logger/timer: it links
``colmap/mvs/images`` -> ``sfm_perspective/images`` and ``colmap/mvs/sparse``
-> ``sfm_perspective/tri_ba`` (``stereo_pipeline.py:397-405``), runs
``reparam_depth`` (``stereo_pipeline.py:408``), creates
``stereo/{depth_maps,normal_maps,consistency_graphs}``
(``stereo_pipeline.py:411-417``) and writes ``patch-match.cfg`` with
``<img>\\n__auto__, 20\\n`` per image (``stereo_pipeline.py:422-424``) and
``fusion.cfg`` (``stereo_pipeline.py:436-438``).

``parse_patch_match_cfg`` / ``select_sources`` resolve the per-reference
source list the way the dense stage reads ``patch-match.cfg``: an explicit
comma-separated list, ``__all__``, or ``__auto__, N``.  For ``__auto__`` the
sources are the other images ranked by the number of sparse points they share
with the reference, among those whose median triangulation angle over the
shared points reaches ``min_triangulation_angle`` (ties by name).
"""
import os

import numpy as np

from .colmap.read_model import read_model, qvec2rotmat
from .reparam_depth import reparam_depth


def prepare_dense_workspace(work_dir, max_src=20):
    colmap_dir = os.path.join(work_dir, "colmap")
    mvs_dir = os.path.join(colmap_dir, "mvs")
    os.makedirs(mvs_dir, exist_ok=True)
    for link, target in (("images", "sfm_perspective/images"), ("sparse", "sfm_perspective/tri_ba")):
        p = os.path.join(mvs_dir, link)
        if os.path.lexists(p):
            os.unlink(p)
        os.symlink(os.path.relpath(os.path.join(colmap_dir, target), mvs_dir), p)
    reparam_depth(os.path.join(mvs_dir, "sparse"), mvs_dir, camera_model="perspective")
    stereo_dir = os.path.join(mvs_dir, "stereo")
    for sub in ("", "depth_maps", "normal_maps", "consistency_graphs"):
        os.makedirs(os.path.join(stereo_dir, sub), exist_ok=True)
    names = sorted(os.listdir(os.path.join(mvs_dir, "images")))
    with open(os.path.join(stereo_dir, "patch-match.cfg"), "w") as fp:
        for n in names:
            fp.write(n + "\n__auto__, {}\n".format(max_src))
    with open(os.path.join(stereo_dir, "fusion.cfg"), "w") as fp:
        for n in names:
            fp.write(n + "\n")
    return mvs_dir


def parse_patch_match_cfg(path):
    """List of (ref_name, spec) where spec is a list of names, ('__all__',)
    or ('__auto__', N)."""
    lines = [ln.strip() for ln in open(path) if ln.strip()]
    out = []
    for i in range(0, len(lines) - 1, 2):
        ref, spec = lines[i], lines[i + 1]
        if spec.startswith("__auto__"):
            n = int(spec.split(",")[1]) if "," in spec else 20
            out.append((ref, ("__auto__", n)))
        elif spec == "__all__":
            out.append((ref, ("__all__",)))
        else:
            out.append((ref, [s.strip() for s in spec.split(",") if s.strip()]))
    return out


def camera_centers(images):
    out = {}
    for im in images.values():
        R = qvec2rotmat(im.qvec)
        out[im.name] = -R.T @ np.asarray(im.tvec, dtype=np.float64)
    return out


def select_sources(mvs_dir, min_triangulation_angle=5.0):
    """Resolve ``patch-match.cfg`` into {ref_name: [src_name, ...]}."""
    _, images, points3D = read_model(os.path.join(mvs_dir, "sparse"), ext=".txt")
    by_name = {im.name: iid for iid, im in images.items()}
    cfg = parse_patch_match_cfg(os.path.join(mvs_dir, "stereo", "patch-match.cfg"))
    order = sorted(images)
    row = {iid: k for k, iid in enumerate(order)}
    pts = list(points3D.values())
    X = np.array([p.xyz for p in pts], dtype=np.float64).reshape(-1, 3)
    vis = np.zeros((len(order), len(pts)), dtype=bool)
    for j, p in enumerate(pts):
        for i in p.image_ids:
            vis[row[int(i)], j] = True
    centers = camera_centers(images)
    C = np.array([centers[images[iid].name] for iid in order])
    rays = C[:, None, :] - X[None, :, :]
    rays /= np.linalg.norm(rays, axis=-1, keepdims=True)
    out = {}
    for ref, spec in cfg:
        if ref not in by_name:
            continue
        a = row[by_name[ref]]
        others = [n for n in sorted(by_name) if n != ref]
        if isinstance(spec, list):
            out[ref] = [s for s in spec if s in by_name and s != ref]
        elif spec[0] == "__all__":
            out[ref] = others
        else:
            cand = []
            for n in others:
                b = row[by_name[n]]
                shared = vis[a] & vis[b]
                if not shared.any():
                    continue
                c = np.clip(np.sum(rays[a, shared] * rays[b, shared], axis=1), -1.0, 1.0)
                if np.median(np.degrees(np.arccos(c))) >= min_triangulation_angle:
                    cand.append((-int(shared.sum()), n))
            cand.sort()
            out[ref] = [n for _, n in cand[:spec[1]]]
    return out
