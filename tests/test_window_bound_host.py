"""The per-(tile, source) interior bound the half-step kernel uses to skip
per-pixel window bounds work, restated on the host (FP64) and checked by
brute force: every tap of every window a tile can take, for hypotheses with
depth in [d_lo, d_hi] and plane slopes |a|, |b| <= amax, projects into the
box spanned by the 8 corner projections of the widened (u, v, d) box the
bound uses; when the bound says "inner", those taps lie inside the source.

Tap of pixel (u, v) (tile-centre relative) at window offset (dx, dy) under
the plane (d, a, b): A (u + dx, v + dy, 1, d + a dx + b dy), A the tile-centred
3 x 4 map into anchor-relative source pixels.
"""
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from test_oracle_pm_search import _satellite_P4  # noqa: E402


def _tile_map(P4r, P4s, uc, vc, dmid, W, H):
    """Tile-centred, anchor-relative map (the kernel's setup in FP64)."""
    M = (P4s[:3] @ np.linalg.inv(P4r)).reshape(-1)
    q = M.reshape(3, 4) @ np.array([uc, vc, 1.0, dmid])
    ax = int(np.rint(min(max(q[0] / q[2], 0.0), W - 1)))
    ay = int(np.rint(min(max(q[1] / q[2], 0.0), H - 1)))
    A = np.zeros(12)
    for r in range(3):
        A[4 * r + 0] = M[4 * r + 0]
        A[4 * r + 1] = M[4 * r + 1]
        A[4 * r + 2] = M[4 * r + 0] * uc + M[4 * r + 1] * vc + M[4 * r + 2]
        A[4 * r + 3] = M[4 * r + 3]
    for k in range(4):
        A[k] -= ax * A[8 + k]
        A[4 + k] -= ay * A[8 + k]
    return A, ax, ay


def _corner_box(A, hu, hv, r, dlo, dhi, amax):
    d0, d1 = dlo - 2 * amax * r, dhi + 2 * amax * r
    pts, zmin = [], math.inf
    for k in range(8):
        u = (hu + r) * (1 if k & 1 else -1)
        v = (hv + r) * (1 if k & 2 else -1)
        d = d1 if k & 4 else d0
        q = A.reshape(3, 4) @ np.array([u, v, 1.0, d])
        if not q[2] > 0:
            return None
        pts.append(q[:2] / q[2])
        zmin = min(zmin, q[2])
    pts = np.array(pts)
    return pts.min(0), pts.max(0), zmin


def test_taps_lie_in_the_corner_box():
    rng = np.random.default_rng(4)
    W = H = 512
    r, amax = 5, 1.01 * math.tan(math.radians(80.0))
    checked = 0
    for geom in [(12.0, 0.0, 24.0, 120.0), (5.0, 45.0, 30.0, 200.0), (20.0, 300.0, 8.0, 60.0)]:
        P4r = _satellite_P4(geom[0], geom[1], 1)
        P4s = _satellite_P4(geom[2], geom[3], 2)
        dmin, dmax = 0.0, 200.0
        dlo, dhi = dmin - 0.5 * (dmax - dmin), dmax + 0.5 * (dmax - dmin)
        for tw in (16, 32, 64):
            th = 512 // tw
            for _ in range(6):
                uc = int(rng.integers(0, W // tw)) * tw + tw // 2
                vc = int(rng.integers(0, H // th)) * th + th // 2
                A, ax, ay = _tile_map(P4r, P4s, uc, vc, 0.5 * (dmin + dmax), W, H)
                box = _corner_box(A, tw // 2, th // 2, r, dlo, dhi, amax)
                assert box is not None
                lo, hi, zmin = box
                n = 4000
                du = rng.integers(-tw // 2, tw // 2, n)
                dv = rng.integers(-th // 2, th // 2, n)
                d = rng.uniform(dlo, dhi, n)
                a = rng.uniform(-amax, amax, n)
                b = rng.uniform(-amax, amax, n)
                dx = rng.integers(-r, r + 1, n)
                dy = rng.integers(-r, r + 1, n)
                X = np.stack([du + dx, dv + dy, np.ones(n), d + a * dx + b * dy])
                q = A.reshape(3, 4) @ X
                assert np.all(q[2] >= zmin * (1 - 1e-12))
                p = q[:2] / q[2]
                assert np.all(p[0] >= lo[0] - 1e-6) and np.all(p[0] <= hi[0] + 1e-6)
                assert np.all(p[1] >= lo[1] - 1e-6) and np.all(p[1] <= hi[1] + 1e-6)
                # the per-tap reciprocal's relative change across a window is
                # bounded by r (|A8| + |A9| + 2 amax |A11|) / zmin
                dz = np.abs(A[8] * dx + A[9] * dy + A[11] * (a * dx + b * dy)) / q[2]
                assert np.all(dz <= r * (abs(A[8]) + abs(A[9]) + 2 * amax * abs(A[11])) / zmin + 1e-15)
                checked += n
    assert checked > 100000


def test_a_window_that_leaves_the_image_is_not_inner():
    """A tile at the image corner: its box reaches outside the source, so the
    bound must not call it inner (the guard of 1.5 px included)."""
    P4r = _satellite_P4(12.0, 0.0, 1)
    P4s = _satellite_P4(24.0, 120.0, 2)
    W = H = 512
    A, ax, ay = _tile_map(P4r, P4s, 16, 8, 100.0, W, H)
    lo, hi, _ = _corner_box(A, 16, 8, 5, -100.0, 300.0, 1.01 * math.tan(math.radians(80.0)))
    g = 1.5
    inside = lo[0] >= -ax + g and hi[0] <= W - 1 - ax - g and lo[1] >= -ay + g and hi[1] <= H - 1 - ay - g
    assert not inside
