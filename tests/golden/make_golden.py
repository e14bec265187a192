"""Writes tests/golden/fusion_golden.json: small fusion cases whose expected
values are derived by hand below (closed-form reasoning on the semantics of
lib/proj_to_grid.py:41-81, produce_dsm.py:58 / aggregate_2p5d.py:81 and
aggregate_2p5d.py:66-81).  Nothing here runs the reference or the oracle.

    python tests/golden/make_golden.py
"""
import json
import os

NAN = None   # JSON null -> NaN

cases = {"proj_to_grid": [], "median_blur3": [], "fuse_cell": []}

# Grid xoff=0, yoff=0, 1 m cells, 3x3.  row = floor((0 - N)/1), col = floor(E/1).
#  (0.5,-0.5,5) and (0.2,-0.3,7) -> cell (0,0), nanmax 7
#  (2.9,-2.1,3) -> cell (2,2)
#  (3.1,-0.5,9) col 3 and (-0.1,-0.5,9) col -1 -> dropped
# Hole fill from the unfilled raster [[7,.,.],[.,.,.],[.,.,3]], in-bounds 3x3:
#  (0,1)->7  (0,2)->none  (1,0)->7  (1,1)->median(7,3)=5  (1,2)->3  (2,0)->none
#  (2,1)->3
cases["proj_to_grid"].append(dict(
    name="nanmax_bounds_holefill",
    points=[[0.5, -0.5, 5.0], [0.2, -0.3, 7.0], [2.9, -2.1, 3.0], [3.1, -0.5, 9.0], [-0.1, -0.5, 9.0]],
    grid=[0.0, 0.0, 1.0, 1.0, 3, 3],
    expected=[[7.0, 7.0, NAN], [7.0, 5.0, 3.0], [NAN, 3.0, 3.0]]))
# Grid xoff=10, yoff=20, 0.5 m cells, 2x2.
#  (10.6,19.4,1.5): row floor(0.6/0.5)=1, col floor(0.6/0.5)=1
#  (10.0,20.0,2.0): row 0, col 0;  (11.0,19.0,8.0): col 2 -> dropped
#  fill (0,1) and (1,0): median(2.0,1.5)=1.75
cases["proj_to_grid"].append(dict(
    name="half_metre_cells_edge",
    points=[[10.6, 19.4, 1.5], [10.0, 20.0, 2.0], [11.0, 19.0, 8.0]],
    grid=[10.0, 20.0, 0.5, 0.5, 2, 2],
    expected=[[2.0, 1.75], [1.75, 1.5]]))
# Empty input: every cell stays NaN.
cases["proj_to_grid"].append(dict(name="no_points", points=[], grid=[0.0, 0.0, 1.0, 1.0, 2, 2],
                                  expected=[[NAN, NAN], [NAN, NAN]]))

# 3x3 median, replicated border: a single spike is removed everywhere
# (every window holds at most one 9 among nine values).
cases["median_blur3"].append(dict(name="spike", input=[[1, 1, 1], [1, 9, 1], [1, 1, 1]],
                                  expected=[[1, 1, 1], [1, 1, 1], [1, 1, 1]]))
# One row [1,2,NaN,4]; rows replicate, so each window is its 3 columns x3:
#  c0: {1,1,2}x3 -> 1;  c1: {1,2}x3 (6 values) -> (1+2)/2 = 1.5
#  c2: {2,4}x3 -> 3;    c3: {4,4}x3 -> 4
cases["median_blur3"].append(dict(name="nan_row", input=[[1, 2, NAN, 4]], expected=[[1, 1.5, 3, 4]]))
cases["median_blur3"].append(dict(name="all_nan", input=[[NAN, NAN], [NAN, NAN]],
                                  expected=[[NAN, NAN], [NAN, NAN]]))

# Cross-view fusion of one cell (then the 3x3 median of a 1x1 raster = itself).
#  [1,2,3,100]: med 2.5; |x-med| = 1.5,0.5,0.5,97.5 -> MAD (0.5+1.5)/2 = 1;
#               keep 2,3 -> mean 2.5
cases["fuse_cell"].append(dict(name="mad_outlier", values=[1, 2, 3, 100], expected=2.5))
#  two measurements (<= 2) -> rejected
cases["fuse_cell"].append(dict(name="two_measurements", values=[5, NAN, 7], expected=None))
#  all equal: MAD 0, nothing is > 0 away -> 4
cases["fuse_cell"].append(dict(name="all_equal", values=[4, 4, 4], expected=4.0))
#  [1,2,4]: med 2; dev 1,0,2 -> MAD 1; keep 1,2 -> 1.5
cases["fuse_cell"].append(dict(name="odd_mad", values=[1, 2, 4], expected=1.5))
#  [10,11,NaN,13,40]: med (11+13)/2 = 12; dev 2,1,1,28 -> MAD 1.5; keep 11,13 -> 12
cases["fuse_cell"].append(dict(name="nan_skipped", values=[10, 11, NAN, 13, 40], expected=12.0))
#  min_measurements 2: [5,NaN,7] -> med 6, MAD 1, keep both -> 6
cases["fuse_cell"].append(dict(name="min_measurements_2", values=[5, NAN, 7], min_measurements=2,
                               expected=6.0))


def _nan(v):
    if isinstance(v, list):
        return [_nan(x) for x in v]
    return v


if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fusion_golden.json")
    with open(out, "w") as fp:
        json.dump(cases, fp, indent=1)
    print(out)
