"""ctypes wrapper and FP64 geometry setup for the PatchMatch oracle.

TEST INFRASTRUCTURE (the oracle): imported only by tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg, as the checker.

``view_geometry`` derives, in numpy FP64 and independently of the product's
C++ host code, everything the per-pixel semantics need from the 4x4
plane-plus-parallax matrices ``P4 = [K[R|t]; depth_min*(0,0,1,-z0)]``
(reference reparam_depth.py:136-141):

* ``M_s = P4_s[:3] inv(P4_ref)``: ref projective point (x, y, 1, d) -> source
  pixel (homogeneous), the inverse projection of aggregate_2p5d_util.py:86-90
  followed by the source projection;
* ``Mb_s = P4_ref[:3] inv(P4_s)``: the backward map for the forward-backward
  check;
* ``X0``: the scene point at the reference image centre and mid depth; the
  triangulation angles, source viewing directions and the Jacobian
  ``J = d(x, y, d)/dX`` of the reference's projective map are taken there.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc_pm.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        L.orc_photo_cost.restype = ctypes.c_double
        L.orc_photo_cost.argtypes = [P(OrcView), P(OrcParams), ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double, P(ctypes.c_double), ctypes.c_int]
        L.orc_tap_coord.restype = ctypes.c_int
        L.orc_tap_coord.argtypes = [P(OrcView), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double, P(ctypes.c_double), ctypes.c_int,
                                    ctypes.c_int, P(ctypes.c_double)]
        L.orc_draw.restype = None
        L.orc_draw.argtypes = [P(OrcView), P(OrcParams), ctypes.c_uint32, ctypes.c_uint32,
                               P(ctypes.c_double), P(ctypes.c_int32)]
        L.orc_random_normal.restype = None
        L.orc_random_normal.argtypes = [P(OrcParams), ctypes.c_double, ctypes.c_double,
                                        P(ctypes.c_double)]
        L.orc_perturb_normal.restype = None
        L.orc_perturb_normal.argtypes = [P(OrcParams), P(ctypes.c_double), ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, P(ctypes.c_double)]
        L.orc_geo_error.restype = ctypes.c_double
        L.orc_geo_error.argtypes = [P(OrcView), P(OrcParams), ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_int]
        _LIB = L
    return _LIB


class OrcParams(ctypes.Structure):
    _fields_ = [("window_radius", ctypes.c_int32), ("window_step", ctypes.c_int32),
                ("sigma_spatial", ctypes.c_double), ("sigma_color", ctypes.c_double),
                ("ncc_sigma", ctypes.c_double), ("num_samples", ctypes.c_int32),
                ("min_tri_angle", ctypes.c_double), ("incident_angle_sigma", ctypes.c_double),
                ("geom_regularizer", ctypes.c_double), ("geom_max_cost", ctypes.c_double),
                ("filter_min_ncc", ctypes.c_double), ("filter_min_tri_angle", ctypes.c_double),
                ("filter_geom_max_cost", ctypes.c_double),
                ("filter_min_num_consistent", ctypes.c_int32),
                ("depth_min", ctypes.c_double), ("depth_max", ctypes.c_double),
                ("perturbation", ctypes.c_double), ("normal_perturbation", ctypes.c_double),
                ("max_normal_angle", ctypes.c_double), ("perturb_decay", ctypes.c_double),
                ("n_far", ctypes.c_int32), ("seed", ctypes.c_uint64), ("draw_group", ctypes.c_int32)]


class OrcView(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("ref", ctypes.POINTER(ctypes.c_uint8)), ("n_src", ctypes.c_int32),
                ("src", ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8))),
                ("src_w", ctypes.POINTER(ctypes.c_int32)), ("src_h", ctypes.POINTER(ctypes.c_int32)),
                ("M", ctypes.POINTER(ctypes.c_double)), ("Mb", ctypes.POINTER(ctypes.c_double)),
                ("tri", ctypes.POINTER(ctypes.c_double)), ("g", ctypes.POINTER(ctypes.c_double)),
                ("J", ctypes.POINTER(ctypes.c_double)),
                ("src_depth", ctypes.POINTER(ctypes.POINTER(ctypes.c_float))),
                ("view_id", ctypes.c_uint32)]


def default_params(**kw):
    """Parameter values of the reference's photometric call
    (colmap_mvs_commands.py:40-48) and consistency call
    (colmap_mvs_commands.py:60-75), plus this project's schedule constants."""
    p = dict(window_radius=5, window_step=1, sigma_spatial=-1.0, sigma_color=0.2,
             ncc_sigma=0.6, num_samples=15, min_triangulation_angle=5.0,
             incident_angle_sigma=0.9, geom_consistency_regularizer=0.3,
             geom_consistency_max_cost=3.0, filter_min_ncc=0.1,
             filter_min_triangulation_angle=4.99, filter_geom_consistency_max_cost=1.0,
             filter_min_num_consistent=2, depth_min=0.0, depth_max=1.0, perturbation=0.05,
             normal_perturbation=0.3, max_normal_angle=80.0, perturb_decay=0.5, n_far=6,
             seed=0x5A7E11173D, draw_group=0)
    p.update(kw)
    return p


def _draw_group_width(p):
    """The oracle's draw key: 0 (per pixel) or the width of the pixel groups
    sharing the draws' uniforms -- the GPU's draw_group = 1 groups the pixels
    of one wave of its half-step kernel, tile_width x (128 / tile_width)."""
    return int(p.get("tile_width", 32)) if p.get("draw_group", 0) else 0


def to_orc_params(p):
    ss = p["sigma_spatial"] if p["sigma_spatial"] > 0 else float(p["window_radius"])
    return OrcParams(p["window_radius"], p["window_step"], ss, p["sigma_color"], p["ncc_sigma"],
                     p["num_samples"], math.radians(p["min_triangulation_angle"]),
                     p["incident_angle_sigma"], p["geom_consistency_regularizer"],
                     p["geom_consistency_max_cost"], p["filter_min_ncc"],
                     math.radians(p["filter_min_triangulation_angle"]),
                     p["filter_geom_consistency_max_cost"], p["filter_min_num_consistent"],
                     p["depth_min"], p["depth_max"], p["perturbation"], p["normal_perturbation"],
                     math.radians(p["max_normal_angle"]), p["perturb_decay"], p["n_far"],
                     p["seed"], _draw_group_width(p))


def camera_center(P4):
    P = P4[:3]
    return -np.linalg.solve(P[:, :3], P[:, 3])


def rq_rotation(P4):
    """R of P[:3,:3] = K R (K upper-triangular, positive diagonal)."""
    A = P4[:3, :3]
    r3 = A[2] / np.linalg.norm(A[2])
    r2 = A[1] - (A[1] @ r3) * r3
    r2 /= np.linalg.norm(r2)
    r1 = np.cross(r2, r3)
    if r1 @ A[0] < 0:
        r1 = -r1
    return np.stack([r1, r2, r3])


def view_geometry(P4_ref, width, height, P4_srcs, depth_range):
    P4_ref = np.asarray(P4_ref, dtype=np.float64)
    inv_r = np.linalg.inv(P4_ref)
    dmid = 0.5 * (depth_range[0] + depth_range[1])
    X0h = inv_r @ np.array([(width - 1) / 2.0, (height - 1) / 2.0, 1.0, dmid])
    X0 = X0h[:3] / X0h[3]
    y = P4_ref @ np.append(X0, 1.0)
    J = np.stack([(P4_ref[i, :3] * y[2] - y[i] * P4_ref[2, :3]) / (y[2] * y[2]) for i in (0, 1, 3)])
    J = J / np.abs(J).max()
    Cr = camera_center(P4_ref)
    M, Mb, tri, g = [], [], [], []
    for P4s in P4_srcs:
        P4s = np.asarray(P4s, dtype=np.float64)
        m = P4s[:3] @ inv_r
        M.append(m / np.abs(m).max())
        mb = P4_ref[:3] @ np.linalg.inv(P4s)
        Mb.append(mb / np.abs(mb).max())
        Cs = camera_center(P4s)
        a = Cr - X0
        b = Cs - X0
        tri.append(math.acos(np.clip(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)), -1, 1)))
        g.append(J @ (b / np.linalg.norm(b)))
    R = rq_rotation(P4_ref)
    return dict(M=np.array(M).reshape(-1, 12), Mb=np.array(Mb).reshape(-1, 12),
                tri=np.array(tri), g=np.array(g).reshape(-1, 3), J=J, X0=X0, R=R,
                N_out=R @ J.T)


def normals_to_camera(normal_uvd, N_out):
    """Unit normals in the reference camera frame, facing the camera (z < 0)."""
    n = normal_uvd.reshape(-1, 3) @ N_out.T
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    n = np.where(ln > 0, n / np.where(ln > 0, ln, 1), 0)
    n = np.where(n[:, 2:3] > 0, -n, n)
    return n.reshape(normal_uvd.shape)


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleView:
    """One reference view with its sources, ready for the C oracle."""

    def __init__(self, ref_img, src_imgs, P4_ref, P4_srcs, depth_range, view_id=0,
                 src_depths=None, params=None):
        self.params = default_params(depth_min=float(depth_range[0]), depth_max=float(depth_range[1]))
        if params:
            self.params.update(params)
        self.ref = np.ascontiguousarray(ref_img, dtype=np.uint8)
        self.srcs = [np.ascontiguousarray(s, dtype=np.uint8) for s in src_imgs]
        self.H, self.W = self.ref.shape
        self.S = len(self.srcs)
        self.geo = view_geometry(P4_ref, self.W, self.H, P4_srcs, depth_range)
        self._M = np.ascontiguousarray(self.geo["M"])
        self._Mb = np.ascontiguousarray(self.geo["Mb"])
        self._tri = np.ascontiguousarray(self.geo["tri"])
        self._g = np.ascontiguousarray(self.geo["g"])
        self._J = np.ascontiguousarray(self.geo["J"].reshape(-1))
        self._sw = np.array([s.shape[1] for s in self.srcs], dtype=np.int32)
        self._sh = np.array([s.shape[0] for s in self.srcs], dtype=np.int32)
        self._srcp = (ctypes.POINTER(ctypes.c_uint8) * self.S)(
            *[s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)) for s in self.srcs])
        self.set_source_depths(src_depths)
        self.view_id = view_id

    def set_source_depths(self, src_depths):
        if src_depths is None:
            self._sd = None
            self._sdp = None
        else:
            self._sd = [np.ascontiguousarray(d, dtype=np.float32) for d in src_depths]
            self._sdp = (ctypes.POINTER(ctypes.c_float) * self.S)(
                *[d.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for d in self._sd])

    def _view(self):
        v = OrcView()
        v.width, v.height = self.W, self.H
        v.ref = self.ref.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        v.n_src = self.S
        v.src = self._srcp
        v.src_w = self._sw.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        v.src_h = self._sh.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        v.M, v.Mb = _dptr(self._M), _dptr(self._Mb)
        v.tri, v.g, v.J = _dptr(self._tri), _dptr(self._g), _dptr(self._J)
        v.src_depth = self._sdp if self._sdp is not None else ctypes.POINTER(ctypes.POINTER(ctypes.c_float))()
        v.view_id = self.view_id
        return v

    def _p(self):
        return to_orc_params(self.params)

    def new_state(self):
        HW = self.H * self.W
        return dict(depth=np.zeros(HW), normal=np.zeros(3 * HW), cost=np.zeros(HW),
                    prob=np.zeros(self.S * HW))

    def init(self, state, pass_=0):
        v, p = self._view(), self._p()
        lib().orc_init(ctypes.byref(v), ctypes.byref(p), pass_, _dptr(state["depth"]),
                       _dptr(state["normal"]), _dptr(state["cost"]), _dptr(state["prob"]))

    def reset(self, state):
        v = self._view()
        lib().orc_reset_state(ctypes.byref(v), _dptr(state["cost"]), _dptr(state["prob"]))

    def halfstep(self, state, pass_, it, color, diagnostics=False):
        v, p = self._view(), self._p()
        HW = self.H * self.W
        win = np.zeros(HW, dtype=np.int32) if diagnostics else None
        cand = np.zeros(HW * 8) if diagnostics else None
        lib().orc_halfstep(ctypes.byref(v), ctypes.byref(p), pass_, it, color,
                           _dptr(state["depth"]), _dptr(state["normal"]), _dptr(state["cost"]),
                           _dptr(state["prob"]),
                           win.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if diagnostics else None,
                           _dptr(cand) if diagnostics else None)
        return (win, cand.reshape(HW, 8)) if diagnostics else None

    def filter(self, state, pass_):
        v, p = self._view(), self._p()
        mask = np.zeros(self.H * self.W, dtype=np.uint64)
        lib().orc_filter(ctypes.byref(v), ctypes.byref(p), pass_, _dptr(state["depth"]),
                         _dptr(state["normal"]), mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        return mask

    def cost_map(self, depth, normal, pass_=0):
        v, p = self._view(), self._p()
        d = np.ascontiguousarray(depth, dtype=np.float64).reshape(-1)
        n = np.ascontiguousarray(normal, dtype=np.float64).reshape(-1)
        out = np.zeros(self.S * self.H * self.W)
        lib().orc_cost_map(ctypes.byref(v), ctypes.byref(p), pass_, _dptr(d), _dptr(n), _dptr(out))
        return out.reshape(self.S, self.H, self.W)

    def tap_coord(self, s, x, y, d, n, dx, dy):
        """Source pixel (px, py) of window tap (dx, dy) of pixel (x, y) under the
        hypothesis (d, n) against source s, and whether its homogeneous depth is
        positive."""
        v = self._view()
        nn = np.ascontiguousarray(n, dtype=np.float64)
        out = np.zeros(2)
        ok = lib().orc_tap_coord(ctypes.byref(v), s, x, y, d, _dptr(nn), dx, dy, _dptr(out))
        return out[0], out[1], bool(ok)

    def draw(self, pixel, code, q):
        """Draw counts per source of the Monte-Carlo view selection of one pixel
        (random stream of (pixel, step code)) for the weights q."""
        v, p = self._view(), self._p()
        qq = np.ascontiguousarray(q, dtype=np.float64)
        assert qq.shape == (self.S,)
        cnt = np.zeros(self.S, dtype=np.int32)
        lib().orc_draw(ctypes.byref(v), ctypes.byref(p), pixel, code, _dptr(qq),
                       cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return cnt

    def photo_cost(self, x, y, d, n, s):
        v, p = self._view(), self._p()
        nn = np.ascontiguousarray(n, dtype=np.float64)
        return lib().orc_photo_cost(ctypes.byref(v), ctypes.byref(p), x, y, d, _dptr(nn), s)

    def solve(self, pass_=0, num_iterations=5, state=None, do_filter=False):
        """Full pass; returns the state dict (depth/normal/cost/prob, mask)."""
        if state is None:
            state = self.new_state()
            self.init(state, pass_)
        else:
            self.reset(state)
        for it in range(num_iterations):
            for color in (0, 1):
                self.halfstep(state, pass_, it, color)
        if do_filter:
            state["mask"] = self.filter(state, pass_)
        return state


def step_code(pass_, it, color):
    """Counter word of the random stream of one half-step."""
    return (pass_ << 24) | (it << 1) | color


def random_normal(params, u1, u2):
    """Random hypothesis normal of the oracle for the uniforms (u1, u2)."""
    p = to_orc_params(default_params(**params))
    n = np.zeros(3)
    lib().orc_random_normal(ctypes.byref(p), u1, u2, _dptr(n))
    return n


def perturb_normal(params, n, amax, u1, u2):
    """Perturbed normal of the oracle: n turned by amax * u1 towards the
    azimuth 2 pi u2 (n itself when the result leaves the cone)."""
    p = to_orc_params(default_params(**params))
    nn = np.ascontiguousarray(n, dtype=np.float64)
    out = np.zeros(3)
    lib().orc_perturb_normal(ctypes.byref(p), _dptr(nn), amax, u1, u2, _dptr(out))
    return out


def philox(ctr, key):
    """Philox4x32-10 block (for known-answer tests)."""
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return tuple(o)
