"""Drop-in for the reference's aggregate_2p5d.py.

``run_fuse(work_dir, max_processes=-1)`` (aggregate_2p5d.py:45-89): per-view
DSMs of the geometric depth maps, then per cell the <= 2-measurement
rejection (aggregate_2p5d.py:69-71), MAD outlier removal and mean
(aggregate_2p5d.py:74-78) and a 3x3 median (aggregate_2p5d.py:81), written
to ``mvs_results/aggregate_2p5d/aggregate_2p5d_dsm.tif`` (GeoTIFF, nodata
-10000 as produce_dsm_from_height writes it, produce_dsm.py:71-82; plus an
``.npy`` with NaN = empty).  The whole chain runs on the device in one call;
for a UTM aoi.json the grid is UTM (see aggregate_2p5d_util).

Keyword extensions: ``depth_type`` (the reference hard-codes 'geometric',
aggregate_2p5d.py:48) and ``min_measurements`` (3 reproduces the reference's
``num_measurements <= 2`` rejection).  The PLY/JPG previews
(aggregate_2p5d.py:84-113) are outside this path.
"""
import logging
import os

import numpy as np

from . import runtime
from .aggregate_2p5d_util import convert_depth_maps


def run_fuse(work_dir, max_processes=-1, depth_type="geometric", min_measurements=3):
    dsm_dir = os.path.join(work_dir, "colmap", "mvs", "dsm")
    names, fused, pv = convert_depth_maps(work_dir, dsm_dir, depth_type, max_processes,
                                          min_measurements)
    for name, dsm in zip(names, pv):
        logging.info("dsm %s empty ratio: %s", name, float(np.isnan(dsm).mean()))
    out_dir = os.path.join(work_dir, "mvs_results", "aggregate_2p5d")
    os.makedirs(out_dir, exist_ok=True)
    runtime.write_dsm_tif_aoi(fused, os.path.join(out_dir, "aggregate_2p5d_dsm.tif"),
                              runtime.read_aoi(work_dir))
    np.save(os.path.join(out_dir, "aggregate_2p5d_dsm.npy"), fused)
    logging.info("After aggregation, empty ratio: %s", float(np.isnan(fused).mean()))
    return fused
