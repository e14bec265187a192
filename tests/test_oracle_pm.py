"""Pins of the PatchMatch oracle (oracle/pm_oracle.c) on the CPU.

The reference holds no PatchMatch code or vectors (the adapted COLMAP binary
is external), so the oracle is pinned by:
* Philox4x32-10 known-answer vectors published with Random123 (Salmon et al.,
  SC'11, kat_vectors: all-zero, all-ones and pi-digit inputs);
* closed forms of the weighted NCC: an exactly affine copy of the window
  (gain +1) gives cost 0, the inverted copy (gain -1) cost 2, a flat window
  the maximum cost, and a brute-force evaluation of the definition on a 3x3
  window through a non-trivial homography matches;
* the forward-backward error through a parallax-only geometry has the closed
  form e = k |d - d0|;
* the filter's truth table.
"""
import math

import numpy as np
import pytest

import pm_oracle as po


def test_philox_known_answers():
    assert po.philox((0, 0, 0, 0), (0, 0)) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)
    assert po.philox((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2) == (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)
    assert po.philox((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0)) == \
        (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)


def _plain_view(ref, src, M, Mb=None, src_depth=None, tri=(0.5,), params=None):
    """An OracleView whose geometry is given directly (bypassing P4)."""
    ov = po.OracleView.__new__(po.OracleView)
    p = po.default_params(depth_min=0.0, depth_max=100.0)
    p.update(params or {})
    ov.params = p
    ov.ref = np.ascontiguousarray(ref, dtype=np.uint8)
    ov.srcs = [np.ascontiguousarray(src, dtype=np.uint8)]
    ov.H, ov.W = ov.ref.shape
    ov.S = 1
    ov._M = np.ascontiguousarray(np.asarray(M, dtype=np.float64).reshape(1, 12))
    ov._Mb = np.ascontiguousarray(np.asarray(Mb if Mb is not None else M, dtype=np.float64).reshape(1, 12))
    ov._tri = np.array(tri, dtype=np.float64)
    ov._g = np.array([[0.0, 0.0, 1.0]])
    ov._J = np.eye(3).reshape(-1).copy()
    ov._sw = np.array([src.shape[1]], dtype=np.int32)
    ov._sh = np.array([src.shape[0]], dtype=np.int32)
    import ctypes
    ov._srcp = (ctypes.POINTER(ctypes.c_uint8) * 1)(ov.srcs[0].ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    ov.set_source_depths(None if src_depth is None else [src_depth])
    ov.view_id = 0
    ov.geo = {}
    return ov


IDENT = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]]


def _texture(h=40, w=40, seed=0):
    rng = np.random.default_rng(seed)
    return rng.integers(20, 200, (h, w)).astype(np.uint8)


def test_ncc_affine_copy_is_zero_cost():
    ref = _texture()
    ov = _plain_view(ref, ref.astype(np.int32) + 30, IDENT)
    n = np.array([0.0, 0.0, 1.0])
    for (x, y) in [(10, 10), (20, 17), (30, 25)]:
        assert ov.photo_cost(x, y, 50.0, n, 0) == pytest.approx(0.0, abs=1e-12)


def test_ncc_inverted_copy_is_max_cost():
    ref = _texture()
    ov = _plain_view(ref, 255 - ref.astype(np.int32), IDENT)
    c = ov.photo_cost(15, 15, 50.0, np.array([0.0, 0.0, 1.0]), 0)
    assert c == pytest.approx(2.0, abs=1e-12)


def test_ncc_flat_window_is_max_cost():
    ref = np.full((40, 40), 100, np.uint8)
    ov = _plain_view(ref, _texture(), IDENT)
    assert ov.photo_cost(20, 20, 50.0, np.array([0.0, 0.0, 1.0]), 0) == 2.0


def test_ncc_centre_outside_source_is_max_cost():
    ref = _texture()
    M = [[1, 0, 100, 0], [0, 1, 0, 0], [0, 0, 1, 0]]      # shifted 100 px off the image
    ov = _plain_view(ref, ref, M)
    assert ov.photo_cost(20, 20, 50.0, np.array([0.0, 0.0, 1.0]), 0) == 2.0


def _bilinear(img, x, y):
    h, w = img.shape
    x = min(max(x, 0.0), w - 1)
    y = min(max(y, 0.0), h - 1)
    x0, y0 = int(math.floor(x)), int(math.floor(y))
    x1, y1 = min(x0 + 1, w - 1), min(y0 + 1, h - 1)
    fx, fy = x - x0, y - y0
    f = img.astype(np.float64)
    return (1 - fy) * ((1 - fx) * f[y0, x0] + fx * f[y0, x1]) + fy * ((1 - fx) * f[y1, x0] + fx * f[y1, x1])


def test_ncc_brute_force_3x3_through_homography():
    """Definition of the bilaterally weighted NCC evaluated by hand."""
    ref = _texture(30, 30, 1)
    src = _texture(30, 30, 2)
    M = [[0.9, 0.05, 2.0, 0.03], [-0.04, 1.1, 1.0, -0.02], [1e-4, -2e-4, 1.0, 1e-5]]
    r, ss, sc = 1, 1.0, 0.2
    ov = _plain_view(ref, src, M, params=dict(window_radius=r, sigma_spatial=ss, sigma_color=sc))
    x, y, d = 12, 14, 40.0
    n = np.array([0.2, -0.1, 0.97])
    n /= np.linalg.norm(n)
    a, b = -n[0] / n[2], -n[1] / n[2]
    Mm = np.array(M)
    W = Wr = Ws = Wrr = Wss = Wrs = 0.0
    rc = ref[y, x] / 255.0
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            rv = ref[y + dy, x + dx] / 255.0
            q = Mm @ np.array([x + dx, y + dy, 1.0, d + a * dx + b * dy])
            sv = _bilinear(src, q[0] / q[2], q[1] / q[2]) / 255.0
            w = math.exp(-(dx * dx + dy * dy) / (2 * ss * ss) - (rv - rc) ** 2 / (2 * sc * sc))
            W += w; Wr += w * rv; Ws += w * sv
            Wrr += w * rv * rv; Wss += w * sv * sv; Wrs += w * rv * sv
    var_r = Wrr / W - (Wr / W) ** 2
    var_s = Wss / W - (Ws / W) ** 2
    ncc = (Wrs / W - (Wr / W) * (Ws / W)) / math.sqrt(var_r * var_s)
    assert ov.photo_cost(x, y, d, n, 0) == pytest.approx(1.0 - ncc, abs=1e-9)


def test_ncc_stride_two_uses_odd_offsets_only():
    """window_step 2 on radius 2 samples offsets {-2, 0, 2}: changing the odd
    ring of the source must not change the cost."""
    ref = _texture(30, 30, 3)
    src = ref.astype(np.int32).copy()
    ov1 = _plain_view(ref, src, IDENT, params=dict(window_radius=2, window_step=2))
    c1 = ov1.photo_cost(15, 15, 10.0, np.array([0.0, 0.0, 1.0]), 0)
    src2 = src.copy()
    src2[14, 14:17] = 0
    src2[16, 14:17] = 0
    src2[15, 14] = src2[15, 16] = 0
    ov2 = _plain_view(ref, src2, IDENT, params=dict(window_radius=2, window_step=2))
    assert ov2.photo_cost(15, 15, 10.0, np.array([0.0, 0.0, 1.0]), 0) == pytest.approx(c1, abs=1e-12)


def test_geometric_error_closed_form():
    """Parallax-only pair: src x = ref x + k d; back: ref x = src x - k d_src.
    With a constant source depth d0 the round trip misses by k |d - d0|."""
    k, d0 = 0.5, 30.0
    M = [[1, 0, 0, k], [0, 1, 0, 0], [0, 0, 1, 0]]
    Mb = [[1, 0, 0, -k], [0, 1, 0, 0], [0, 0, 1, 0]]
    ref = _texture(40, 60)
    sd = np.full((40, 60), d0, np.float32)
    ov = _plain_view(ref, ref, M, Mb, src_depth=sd)
    for d in (30.0, 32.0, 25.5):
        assert ov_geo(ov, 10, 10, d) == pytest.approx(k * abs(d - d0), abs=1e-12)
    # invalid source depth -> geom_max
    sd[:] = 0.0
    ov.set_source_depths([sd])
    assert ov_geo(ov, 10, 10, 30.0) == ov.params["geom_consistency_max_cost"]


def ov_geo(ov, x, y, d):
    import ctypes
    v, p = ov._view(), ov._p()
    return po.lib().orc_geo_error(ctypes.byref(v), ctypes.byref(p), x, y, d, 0)


def test_filter_truth_table():
    """Consistent iff NCC >= min_ncc and tri >= min angle (photometric pass)."""
    ref = _texture()
    good = _plain_view(ref, ref, IDENT, tri=(math.radians(10),),
                       params=dict(filter_min_num_consistent=1))
    st = dict(depth=np.full(1600, 20.0), normal=np.tile([0.0, 0.0, 1.0], 1600))
    mask = good.filter(st, 0)
    assert np.all(mask.reshape(40, 40)[5:-5, 5:-5] == 1)
    low_tri = _plain_view(ref, ref, IDENT, tri=(math.radians(2),), params=dict(filter_min_num_consistent=1))
    st = dict(depth=np.full(1600, 20.0), normal=np.tile([0.0, 0.0, 1.0], 1600))
    assert np.all(low_tri.filter(st, 0) == 0)
    assert np.all(st["depth"] == 0.0)
    inv = _plain_view(ref, 255 - ref.astype(np.int32), IDENT, tri=(math.radians(10),),
                      params=dict(filter_min_num_consistent=1))
    st = dict(depth=np.full(1600, 20.0), normal=np.tile([0.0, 0.0, 1.0], 1600))
    assert np.all(inv.filter(st, 0).reshape(40, 40)[5:-5, 5:-5] == 0)


def test_halfstep_keeps_true_plane_and_only_updates_its_colour():
    """On an exactly matched pair every pixel already holds a zero-cost
    hypothesis; a half-step may only keep a zero-cost one, and pixels of the
    other colour are untouched."""
    ref = _texture(24, 24, 5)
    ov = _plain_view(ref, ref.astype(np.int32) + 5, IDENT, params=dict(window_radius=2))
    st = ov.new_state()
    ov.init(st, 0)
    st["depth"][:] = 40.0
    st["normal"][:] = np.tile([0.0, 0.0, 1.0], 24 * 24)
    before = {k: v.copy() for k, v in st.items()}
    win, cand = ov.halfstep(st, 0, 0, 0, diagnostics=True)
    yy, xx = np.mgrid[0:24, 0:24]
    other = ((xx + yy) % 2 == 1).reshape(-1)
    assert np.array_equal(st["depth"][other], before["depth"][other])
    act = ~other
    assert np.all(st["cost"][act] <= 1e-12)
    assert np.all(cand[act, 0] <= 1e-12)
