"""Drop-in for the reference's produce_dsm.py.

``produce_dsm_from_points(work_dir, points, tif_to_write, jpg_to_write=None)``
(produce_dsm.py:46-67): grid from aoi.json with 0.5 m cells
(produce_dsm.py:41-55), proj_to_grid (device), 3x3 median (device,
produce_dsm.py:58), GeoTIFF with nodata -10000 (produce_dsm.py:59-61).
Returns the grid tuple as the reference does.  ``produce_dsm_from_height``
(produce_dsm.py:71-88) writes a given raster.  The JPG previews
(plot_height_map) are outside this path.
"""
import numpy as np

from . import runtime
from .lib.proj_to_grid import proj_to_grid

e_resolution = 0.5
n_resolution = 0.5


def produce_dsm_from_points(work_dir, points, tif_to_write, jpg_to_write=None):
    aoi = runtime.read_aoi(work_dir)
    ul_e, ul_n = aoi["ul_easting"], aoi["ul_northing"]
    e_size = int(aoi["width"] / e_resolution) + 1
    n_size = int(aoi["height"] / n_resolution) + 1
    dsm = proj_to_grid(points, ul_e, ul_n, e_resolution, n_resolution, e_size, n_size, propagate=True)
    dsm = runtime.context().median_blur3(dsm.astype(np.float32))
    if tif_to_write:
        runtime.write_dsm_tif_aoi(dsm, tif_to_write, aoi, e_resolution)
    return (ul_e, ul_n, e_size, n_size, e_resolution, n_resolution)


def produce_dsm_from_height(work_dir, height, tif_to_write, jpg_to_write=None):
    aoi = runtime.read_aoi(work_dir)
    n_size, e_size = height.shape[:2]
    runtime.write_dsm_tif_aoi(height, tif_to_write, aoi, e_resolution)
    return (aoi["ul_easting"], aoi["ul_northing"], e_size, n_size, e_resolution, n_resolution)
