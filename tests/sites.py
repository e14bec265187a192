"""Synthetic sites shared by the tests (built once per session)."""
import os

import numpy as np
from PIL import Image

from vissatsatellitestereo_amd import synth, workspace
from vissatsatellitestereo_amd import reparam_depth as rd

_CACHE = {}


def tiny_spec():
    """4 views of 160x128: multi-source, geometric pass, ragged (non-square)."""
    return synth.SceneSpec(width=160, height=128, n_views=4, off_nadir=(10.0, 25.0),
                           azimuths=(0.0, 90.0, 180.0, 270.0), n_boxes=6, n_points=800)


def build(name, spec, root, max_src=20, geo=None, save_gt_views=True):
    """geo=(lat, lon) anchors the site in UTM (aoi.json with a real zone).
    Sites of more than 4 Mpx in total are rendered on the GPU when there is
    one (small sites always on the CPU)."""
    key = (name, str(root))
    if key in _CACHE:
        return _CACHE[key]
    wd = os.path.join(str(root), name)
    if not os.path.exists(os.path.join(wd, "colmap", "mvs", "depth_ranges.txt")):
        sizes = spec.sizes or [(spec.width, spec.height)] * spec.n_views
        device = "cpu"
        if sum(w * h for w, h in sizes) > 4 * 2**20:
            import torch
            device = "cuda" if torch.cuda.is_available() else "cpu"
        synth.build_site(spec, wd, geo=geo, device=device, save_gt_views=save_gt_views)
        workspace.prepare_dense_workspace(wd, max_src=max_src)
    mvs = os.path.join(wd, "colmap", "mvs")
    P4 = rd.projection_matrices(mvs)
    names = sorted(P4)
    imgs = {n: np.array(Image.open(os.path.join(mvs, "images", n))) for n in names}
    dr = rd.read_name_values(os.path.join(mvs, "depth_ranges.txt"))
    lr = rd.read_name_values(os.path.join(mvs, "last_rows.txt"))
    site = dict(wd=wd, mvs=mvs, P4=P4, names=names, imgs=imgs, dr=dr, last_rows=lr)
    _CACHE[key] = site
    return site


def gt_depth(site, name):
    """True reparametrised depth of every pixel: depth_min (z - z0) / z_cam."""
    gt = os.path.join(site["wd"], "synthetic_gt")
    hz = np.load(os.path.join(gt, name[:-4] + ".hit_z.npy")).astype(np.float64)
    zc = np.load(os.path.join(gt, name[:-4] + ".zcam.npy"))
    a, b, c, d = site["last_rows"][name]
    return (c * hz + d) / zc
