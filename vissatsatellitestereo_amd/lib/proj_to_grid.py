"""Drop-in for the reference's lib/proj_to_grid.py on the device.

``proj_to_grid(points, xoff, yoff, xresolution, yresolution, xsize, ysize,
propagate=False)`` (lib/proj_to_grid.py:41): row = floor((yoff - N)/xres),
col = floor((E - xoff)/yres) (the reference's pairing,
lib/proj_to_grid.py:42-43), out-of-grid points dropped, per-cell nanmax
(lib/proj_to_grid.py:61), then every empty cell takes np.median of the
in-bounds non-NaN 3x3 neighbours of the unfilled raster
(lib/proj_to_grid.py:64-79).  ``propagate`` is accepted and ignored, as in
the reference.  Returns a (ysize, xsize) float64 raster.
"""
from .. import runtime


def proj_to_grid(points, xoff, yoff, xresolution, yresolution, xsize, ysize, propagate=False):
    grid = (float(xoff), float(yoff), float(xresolution), float(yresolution), int(xsize), int(ysize))
    return runtime.context().proj_to_grid(points, grid)
