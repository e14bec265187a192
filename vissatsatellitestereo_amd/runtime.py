"""Process-wide device context shared by the drop-in modules, plus the
workspace plumbing they have in common.

The reference runs `colmap patch_match_stereo` as a subprocess per call
(lib/run_cmd.py:44), so nothing survives between the photometric call, the
consistency call and fusion except files.  Here the three share one device
context: images and depth/normal maps stay resident in HBM
between calls of the same process, and everything is still written to the
workspace files, so a fresh process resumes from them exactly as the
reference does.
"""
import json
import os
import threading

import numpy as np
from PIL import Image

from .colmap.read_dense import read_array
from .mvs import MvsContext
from . import reparam_depth as rd

_LOCK = threading.Lock()
_CTX = None
_STATE = {}


def context():
    global _CTX
    with _LOCK:
        if _CTX is None:
            budget = int(float(os.environ.get("VSM_HBM_BUDGET_GB", "200")) * 1e9)
            _CTX = MvsContext(device=int(os.environ.get("VSM_DEVICE", "0")), hbm_budget_bytes=budget)
    return _CTX


def reset_context():
    global _CTX
    with _LOCK:
        if _CTX is not None:
            _CTX.close()
        _CTX = None
        _STATE.clear()


class Workspace:
    """Image ids, P4 matrices and residency of one dense workspace."""

    def __init__(self, mvs_dir):
        self.mvs_dir = os.path.abspath(mvs_dir)
        self.P4 = rd.projection_matrices(self.mvs_dir)
        self.names = sorted(self.P4)
        self.ids = {n: i + 1 for i, n in enumerate(self.names)}
        self.loaded = set()          # names uploaded with pixels
        self.geometry = set()        # names registered (pixels or geometry only)
        self.photo_resident = set()  # names whose photometric result is on the device
        self.sizes = {}

    @staticmethod
    def get(mvs_dir):
        key = os.path.abspath(mvs_dir)
        ws = _STATE.get(key)
        if ws is None:
            ws = Workspace(mvs_dir)
            _STATE[key] = ws
        return ws

    def image(self, name):
        return np.array(Image.open(os.path.join(self.mvs_dir, "images", name)).convert("L"))

    def upload(self, ctx, names):
        for n in names:
            if n in self.loaded:
                continue
            img = self.image(n)
            ctx.add_image(self.ids[n], img, self.P4[n])
            self.loaded.add(n)
            self.geometry.add(n)
            self.sizes[n] = img.shape

    def register(self, ctx, name, shape):
        if name in self.geometry:
            return
        ctx.register_geometry(self.ids[name], shape[1], shape[0], self.P4[name])
        self.geometry.add(name)
        self.sizes[name] = shape

    def depth_ranges(self):
        return rd.read_name_values(os.path.join(self.mvs_dir, "depth_ranges.txt"))

    def write_inv_proj_mats(self):
        """inv(P4) per image, 16 values row-major, single-space separated
        (the file aggregate_2p5d_util.py:55-61 parses)."""
        with open(os.path.join(self.mvs_dir, "inv_proj_mats.txt"), "w") as fp:
            for n in self.names:
                inv = np.linalg.inv(self.P4[n])
                fp.write(n + " " + " ".join(repr(float(v)) for v in inv.reshape(-1)) + "\n")

    def read_inv_proj_mats(self):
        out = {}
        with open(os.path.join(self.mvs_dir, "inv_proj_mats.txt")) as fp:
            for line in fp.readlines():
                tmp = line.split(" ")
                out[tmp[0]] = np.array([float(tmp[i]) for i in range(1, 17)]).reshape(4, 4)
        return out

    def depth_path(self, name, kind):
        return os.path.join(self.mvs_dir, "stereo", "depth_maps", f"{name}.{kind}.bin")

    def normal_path(self, name, kind):
        return os.path.join(self.mvs_dir, "stereo", "normal_maps", f"{name}.{kind}.bin")

    def read_depth(self, name, kind):
        return read_array(self.depth_path(name, kind)).astype(np.float32)

    def read_normal(self, name, kind):
        return read_array(self.normal_path(name, kind)).astype(np.float32)


def read_aoi(work_dir):
    with open(os.path.join(work_dir, "aoi.json")) as fp:
        return json.load(fp)


def dsm_grid(aoi, e_resolution=0.5, n_resolution=0.5):
    """Grid of produce_dsm.py:51-55: e_size = int(width/res) + 1, ..."""
    e_size = int(aoi["width"] / e_resolution) + 1
    n_size = int(aoi["height"] / n_resolution) + 1
    return (aoi["ul_easting"], aoi["ul_northing"], e_resolution, n_resolution, e_size, n_size)


def write_dsm_tif_aoi(arr, path, aoi, resolution=0.5, nodata_val=-10000):
    """GeoTIFF of a DSM on the aoi.json grid, nodata -10000 as at
    produce_dsm.py:59-61 and produce_dsm.py:80-82 (lib/dsm_util.py:122-159)."""
    from .lib.dsm_util import write_dsm_tif
    write_dsm_tif(np.asarray(arr, dtype=np.float32), path,
                  (aoi["ul_easting"], aoi["ul_northing"], resolution, resolution),
                  (int(aoi.get("zone_number", 0)), aoi.get("hemisphere", "N")), nodata_val=nodata_val)


def read_tif(path):
    """DSM raster of a GeoTIFF written by write_dsm_tif_aoi (nodata -> NaN)."""
    from .lib.dsm_util import read_dsm_tif
    return read_dsm_tif(path)[0]
