"""Pins of the fusion oracle (oracle/fusion_oracle.py) on the CPU.

* proj_to_grid against a literal loop restatement of lib/proj_to_grid.py:41-81
  on small random inputs (integer cell indices compared exactly);
* hand-computed golden cases (tests/golden/fusion_golden.json, written by
  tests/golden/make_golden.py from closed-form reasoning, not by running the
  reference);
* back-projection round trip X -> P4 -> (col, row, d) -> inv(P4) -> X.
"""
import json
import math
import os

import numpy as np

import fusion_oracle as fo

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fusion_golden.json")


def _loop_proj_to_grid(points, xoff, yoff, xres, yres, xsize, ysize):
    """lib/proj_to_grid.py:41-81 as plain loops."""
    dsm = np.full((ysize, xsize), np.nan)
    for E, N, h in points:
        row = math.floor((yoff - N) / xres)
        col = math.floor((E - xoff) / yres)
        if 0 <= row < ysize and 0 <= col < xsize and not math.isnan(h):
            if math.isnan(dsm[row, col]) or h > dsm[row, col]:
                dsm[row, col] = h
    out = dsm.copy()
    for row in range(ysize):
        for col in range(xsize):
            if not math.isnan(dsm[row, col]):
                continue
            nb = [dsm[j, k] for j in range(row - 1, row + 2) for k in range(col - 1, col + 2)
                  if 0 <= j < ysize and 0 <= k < xsize and not math.isnan(dsm[j, k])]
            if nb:
                out[row, col] = np.median(nb)
    return out


def test_proj_to_grid_matches_loop_restatement():
    rng = np.random.default_rng(0)
    for trial in range(5):
        n = int(rng.integers(0, 60))
        pts = np.column_stack([rng.uniform(-1, 6, n), rng.uniform(-6, 1, n), rng.normal(10, 3, n)])
        a = fo.proj_to_grid(pts, 0.0, 0.0, 0.5, 0.5, 11, 9)
        b = _loop_proj_to_grid(pts, 0.0, 0.0, 0.5, 0.5, 11, 9)
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


def test_golden_cases():
    g = json.load(open(GOLD))
    for case in g["proj_to_grid"]:
        out = fo.proj_to_grid(np.array(case["points"]), *case["grid"])
        exp = np.array(case["expected"], dtype=np.float64)
        assert np.array_equal(np.isnan(out), np.isnan(exp)), case["name"]
        assert np.allclose(out[~np.isnan(out)], exp[~np.isnan(exp)], rtol=0, atol=0), case["name"]
    for case in g["median_blur3"]:
        out = fo.median_blur3(np.array(case["input"], dtype=np.float32))
        exp = np.array(case["expected"], dtype=np.float32)
        assert np.array_equal(np.isnan(out), np.isnan(exp)), case["name"]
        assert np.array_equal(out[~np.isnan(out)], exp[~np.isnan(exp)]), case["name"]
    for case in g["fuse_cell"]:
        stack = np.array(case["values"], dtype=np.float32).reshape(-1, 1, 1)
        out = fo.fuse(stack, case.get("min_measurements", 3))
        if case["expected"] is None:
            assert np.isnan(out[0, 0]), case["name"]
        else:
            assert out[0, 0] == np.float32(case["expected"]), case["name"]


def test_backproject_round_trip():
    rng = np.random.default_rng(1)
    K = np.array([[1.0e6, 3.0, 500.0], [0, 1.0e6, 480.0], [0, 0, 1]])
    th = math.radians(20)
    R = np.array([[1, 0, 0], [0, math.cos(th), -math.sin(th)], [0, math.sin(th), math.cos(th)]]) @ \
        np.diag([1.0, -1.0, -1.0])
    C = np.array([0.0, -2.0e5, 6.0e5])
    t = -R @ C
    z0 = -40.0
    dmin = 6.3e5
    P4 = np.vstack([K @ np.hstack([R, t[:, None]]), dmin * np.array([0, 0, 1.0, -z0])])
    X = np.column_stack([rng.uniform(-50, 50, 200), rng.uniform(-50, 50, 200), rng.uniform(-10, 40, 200)])
    y = (P4 @ np.vstack([X.T, np.ones(200)])).T
    col, row, d = y[:, 0] / y[:, 2], y[:, 1] / y[:, 2], y[:, 3] / y[:, 2]
    inv = np.linalg.inv(P4)
    Xb = [((inv[i, 0] * col + inv[i, 1] * row) + inv[i, 2]) + inv[i, 3] * d for i in range(4)]
    Xr = np.column_stack([Xb[0] / Xb[3], Xb[1] / Xb[3], Xb[2] / Xb[3]])
    assert np.abs(Xr - X).max() < 1e-4
    # reparametrised depth is the scaled height above the reference plane
    zc = (R @ X.T + t[:, None])[2]
    assert np.allclose(d, dmin * (X[:, 2] - z0) / zc, rtol=1e-9)


def test_backproject_skips_invalid_depths():
    inv = np.eye(4)
    depth = np.array([[1.0, 0.0], [-1.0, 2.0]], np.float32)
    pts = fo.backproject(depth, inv)
    assert pts.shape == (2, 3)
    # identity inverse: X = (col, row, 1) / d
    assert np.allclose(pts[0], [0.0, 0.0, 1.0])
    assert np.allclose(pts[1], [0.5, 0.5, 0.5])
