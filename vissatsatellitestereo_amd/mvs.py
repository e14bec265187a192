"""Python face of libvissat_mvs.so: a device context holding images, depth
maps and hypotheses in HBM, with batched solves, serving and DSM fusion.

Every method calls the HIP library through the C ABI; there is no CPU path.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import VsmError, fptr, u32ptr, make_req


class MvsContext:
    """One device context (``vsm_create``).

    Parameters mirror ``PatchMatchStereo.*`` flags (``vsm_params``); the
    defaults are the reference's photometric call
    (``colmap_mvs_commands.py:40-48``).
    """

    def __init__(self, device=0, hbm_budget_bytes=200 * 10**9, **params):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self.lib.vsm_create(int(device), int(hbm_budget_bytes), ctypes.byref(h))
        if rc != 0:
            raise VsmError(rc, "vsm_create failed (is a HIP device present?)")
        self.h = h
        self.sizes = {}
        self.set_params(**params)

    # ------------------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            msg = self.lib.vsm_last_error(self.h)
            raise VsmError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self.lib.vsm_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        p = _lib.Params()
        self._check(self.lib.vsm_get_params(self.h, ctypes.byref(p)))
        return p

    def set_params(self, **kw):
        p = self.params
        for k, v in kw.items():
            if not hasattr(p, k):
                raise KeyError(f"unknown parameter {k}")
            setattr(p, k, v)
        self._check(self.lib.vsm_set_params(self.h, ctypes.byref(p)))

    # ------------------------------------------------------------------
    def add_image(self, image_id, gray, P4):
        P = np.ascontiguousarray(P4, dtype=np.float64).reshape(16)
        if gray is None:
            raise ValueError("use register_geometry for images without pixels")
        g = np.ascontiguousarray(gray, dtype=np.uint8)
        if g.ndim != 2:
            raise ValueError("gray image must be 2-D uint8")
        h, w = g.shape
        self._check(self.lib.vsm_add_image(self.h, int(image_id),
                                           g.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                           w, h, w, P.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        self.sizes[int(image_id)] = (h, w)

    def register_geometry(self, image_id, width, height, P4):
        P = np.ascontiguousarray(P4, dtype=np.float64).reshape(16)
        self._check(self.lib.vsm_add_image(self.h, int(image_id), None, int(width), int(height),
                                           int(width), P.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        self.sizes[int(image_id)] = (int(height), int(width))

    def remove_image(self, image_id):
        self._check(self.lib.vsm_remove_image(self.h, int(image_id)))
        self.sizes.pop(int(image_id), None)

    def set_normal_map(self, image_id, normal, kind=1):
        """Camera-frame unit normals (H x W x 3) that go with the depth map."""
        n = np.ascontiguousarray(normal, dtype=np.float32)
        h, w = n.shape[:2]
        self._check(self.lib.vsm_set_normal_map(self.h, int(image_id), int(kind), fptr(n), w, h))

    def fuse_points(self, image_ids, cam_to_world, depth_kind=1, inv_proj_mats=None, capacity=None,
                    **thresholds):
        """3D fusion (vsm_fuse_points).  Returns (xyz N x 3 FP64, normals N x 3
        float32, grey N uint8)."""
        p = _lib.Fuse3dParams()
        self.lib.vsm_default_fuse3d_params(ctypes.byref(p))
        for k, v in thresholds.items():
            setattr(p, k, v)
        ids = np.ascontiguousarray(image_ids, dtype=np.int32)
        R = np.ascontiguousarray(cam_to_world, dtype=np.float64).reshape(len(ids), 9)
        mats = None
        if inv_proj_mats is not None:
            mats = np.ascontiguousarray(inv_proj_mats, dtype=np.float64).reshape(len(ids), 16)
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None
        cap = int(capacity) if capacity is not None else sum(
            self.sizes[int(i)][0] * self.sizes[int(i)][1] for i in ids) // 2 + 1
        while True:
            xyz = np.zeros((cap, 3), np.float64)
            nrm = np.zeros((cap, 3), np.float32)
            gry = np.zeros(cap, np.uint8)
            cnt = ctypes.c_int64()
            self._check(self.lib.vsm_fuse_points(
                self.h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(ids), int(depth_kind),
                ctypes.byref(p), dp(mats), dp(R), cap, dp(xyz), fptr(nrm),
                gry.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(cnt)))
            if cnt.value <= cap:
                return xyz[:cnt.value], nrm[:cnt.value], gry[:cnt.value]
            cap = cnt.value

    def get_depth_map(self, image_id, kind=1):
        """The depth map (kind 0 photometric, 1 geometric) held on the device."""
        h, w = self.sizes[int(image_id)]
        out = np.zeros((h, w), np.float32)
        self._check(self.lib.vsm_get_depth_map(self.h, int(image_id), int(kind), fptr(out), w, h))
        return out

    def set_depth_map(self, image_id, depth, kind=0):
        d = np.ascontiguousarray(depth, dtype=np.float32)
        h, w = d.shape
        self._check(self.lib.vsm_set_depth_map(self.h, int(image_id), int(kind), fptr(d), w, h))

    # ------------------------------------------------------------------
    def _reqs(self, views):
        n = len(views)
        reqs = (_lib.ViewReq * n)()
        keep = []
        for i, v in enumerate(views):
            r, k = make_req(v["ref"], v["srcs"], v["depth_range"], v.get("init_depth"),
                            v.get("init_normal"))
            reqs[i] = r
            keep.append(k)
            if int(v["ref"]) not in self.sizes:
                raise VsmError(-2, f"unknown reference image {int(v['ref'])}")
        return reqs, keep

    def _outputs(self, views, want):
        outs = (_lib.ViewOut * len(views))()
        results = []
        for i, v in enumerate(views):
            h, w = self.sizes[int(v["ref"])]
            res = {}
            if "depth" in want:
                res["depth"] = np.zeros((h, w), np.float32)
                outs[i].depth = fptr(res["depth"])
            if "normal" in want:
                res["normal"] = np.zeros((h, w, 3), np.float32)
                outs[i].normal = fptr(res["normal"])
            if "cost" in want:
                res["cost"] = np.zeros((h, w), np.float32)
                outs[i].cost = fptr(res["cost"])
            if "consistency" in want:
                res["consistency"] = np.zeros((h, w), np.uint32)
                outs[i].consistency = u32ptr(res["consistency"])
            results.append(res)
        return outs, results

    def solve(self, views, pass_=0, want=("depth", "normal", "cost", "consistency"),
              use_graph=False):
        """views: list of dicts with ref, srcs, depth_range and optionally
        init_depth / init_normal.  Returns one dict of outputs per view."""
        reqs, keep = self._reqs(views)
        outs, results = self._outputs(views, want)
        self._check(self.lib.vsm_solve(self.h, reqs, len(views), int(pass_), outs,
                                       1 if use_graph else 0))
        return results

    def batch(self, views, pass_=0, use_graph=False):
        """A staged solve (vsm_batch_begin): step() runs half-steps, end()
        returns the outputs as solve() does."""
        reqs, keep = self._reqs(views)
        b = ctypes.c_int32()
        self._check(self.lib.vsm_batch_begin(self.h, reqs, len(views), int(pass_),
                                             1 if use_graph else 0, ctypes.byref(b)))
        return SolveBatch(self, b.value, views)

    def eval_cost(self, ref, srcs, depth_range, depth, normal_uvd, pass_=0):
        r, keep = make_req(ref, srcs, depth_range)
        h, w = self.sizes[int(ref)]
        d = np.ascontiguousarray(depth, dtype=np.float32).reshape(h, w)
        nn = np.ascontiguousarray(normal_uvd, dtype=np.float32).reshape(h, w, 3)
        out = np.zeros((len(srcs), h, w), np.float32)
        self._check(self.lib.vsm_eval_cost(self.h, ctypes.byref(r), int(pass_), fptr(d), fptr(nn),
                                           fptr(out)))
        return out

    def eval_cost_staging(self, ref, srcs, depth_range, depth, normal_uvd, stage, repeats=5, pass_=0):
        """vsm_eval_cost_staging: (costs, kernel ms per launch, staged fraction)."""
        r, keep = make_req(ref, srcs, depth_range)
        h, w = self.sizes[int(ref)]
        d = np.ascontiguousarray(depth, dtype=np.float32).reshape(h, w)
        nn = np.ascontiguousarray(normal_uvd, dtype=np.float32).reshape(h, w, 3)
        out = np.zeros((len(srcs), h, w), np.float32)
        ms, frac = ctypes.c_float(), ctypes.c_float()
        self._check(self.lib.vsm_eval_cost_staging(self.h, ctypes.byref(r), int(pass_), fptr(d), fptr(nn),
                                                   int(stage), int(repeats), fptr(out), ctypes.byref(ms),
                                                   ctypes.byref(frac)))
        return out, ms.value, frac.value

    def debug_init(self, ref, srcs, depth_range, pass_=0):
        r, keep = make_req(ref, srcs, depth_range)
        h, w = self.sizes[int(ref)]
        st = dict(depth=np.zeros(h * w, np.float32), normal=np.zeros(3 * h * w, np.float32),
                  cost=np.zeros(h * w, np.float32), prob=np.zeros(len(srcs) * h * w, np.float32))
        self._check(self.lib.vsm_debug_init(self.h, ctypes.byref(r), int(pass_), fptr(st["depth"]),
                                            fptr(st["normal"]), fptr(st["cost"]), fptr(st["prob"])))
        return st

    def debug_halfstep(self, ref, srcs, depth_range, state, pass_, it, color):
        """One half-step on a copy of state (dict of float32 arrays)."""
        r, keep = make_req(ref, srcs, depth_range)
        st = {k: np.ascontiguousarray(v, dtype=np.float32).copy() for k, v in state.items()
              if k in ("depth", "normal", "cost", "prob")}
        self._check(self.lib.vsm_debug_halfstep(self.h, ctypes.byref(r), int(pass_), int(it),
                                                int(color), fptr(st["depth"]), fptr(st["normal"]),
                                                fptr(st["cost"]), fptr(st["prob"])))
        return st

    def debug_filter(self, ref, srcs, depth_range, state, pass_):
        r, keep = make_req(ref, srcs, depth_range)
        st = {k: np.ascontiguousarray(state[k], dtype=np.float32).copy() for k in ("depth", "normal")}
        mask = np.zeros(st["depth"].size, np.uint32)
        self._check(self.lib.vsm_debug_filter(self.h, ctypes.byref(r), int(pass_), fptr(st["depth"]),
                                              fptr(st["normal"]), u32ptr(mask)))
        st["mask"] = mask
        return st

    # ------------------------------------------------------------------
    def fuse_dsm(self, grid, image_ids, depth_kind=1, min_measurements=3, inv_proj_mats=None,
                 per_view=False):
        g = _lib.DsmGrid(*grid) if not isinstance(grid, _lib.DsmGrid) else grid
        ids = np.ascontiguousarray(image_ids, dtype=np.int32)
        out = np.zeros((g.n_size, g.e_size), np.float32)
        pv = np.zeros((len(ids), g.n_size, g.e_size), np.float32) if per_view else None
        mats = None
        if inv_proj_mats is not None:
            mats = np.ascontiguousarray(inv_proj_mats, dtype=np.float64).reshape(len(ids), 16)
        self._check(self.lib.vsm_fuse_dsm(
            self.h, ctypes.byref(g), ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(ids),
            int(depth_kind), int(min_measurements),
            mats.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if mats is not None else None,
            fptr(out), fptr(pv)))
        return (out, pv) if per_view else out

    def proj_to_grid(self, points, grid):
        g = _lib.DsmGrid(*grid) if not isinstance(grid, _lib.DsmGrid) else grid
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((g.n_size, g.e_size), np.float64)
        self._check(self.lib.vsm_proj_to_grid(
            self.h, pts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if len(pts) else None,
            len(pts), ctypes.byref(g), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def set_geo_frame(self, geo):
        """geo: None (local ENU DSM grid), a GeoFrame, or the tuple
        (lat0_deg, lon0_deg, alt0, zone_number, south)."""
        if geo is None:
            self._check(self.lib.vsm_set_geo_frame(self.h, None))
            return None
        f = geo if isinstance(geo, _lib.GeoFrame) else _lib.geo_frame(*geo)
        self._check(self.lib.vsm_set_geo_frame(self.h, ctypes.byref(f)))
        return f

    def enu_to_utm(self, enu):
        """n x 3 local ENU -> n x 3 (easting, northing, altitude) on the device."""
        pts = np.ascontiguousarray(enu, dtype=np.float64).reshape(-1, 3)
        out = np.zeros_like(pts)
        if len(pts):
            self._check(self.lib.vsm_enu_to_utm(
                self.h, pts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(pts),
                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def median_blur3(self, img):
        a = np.ascontiguousarray(img, dtype=np.float32)
        out = np.zeros_like(a)
        n_size, e_size = a.shape
        self._check(self.lib.vsm_median_blur3(self.h, fptr(a), e_size, n_size, fptr(out)))
        return out

    # ------------------------------------------------------------------
    def submit(self, ref, srcs, depth_range, pass_=0, filter=None):
        """Queue a served solve (vsm_submit_pass); filter None follows the
        context's filter parameter."""
        r, keep = make_req(ref, srcs, depth_range)
        t = ctypes.c_uint64()
        f = -1 if filter is None else (1 if filter else 0)
        self._check(self.lib.vsm_submit_pass(self.h, ctypes.byref(r), int(pass_), f, ctypes.byref(t)))
        return t.value

    def set_hbm_budget(self, nbytes):
        self._check(self.lib.vsm_set_hbm_budget(self.h, int(nbytes)))

    def serve_step(self):
        na, nq = ctypes.c_int32(), ctypes.c_int32()
        self._check(self.lib.vsm_serve_step(self.h, ctypes.byref(na), ctypes.byref(nq)))
        return na.value, nq.value

    def poll(self, ticket, ref, want=("depth",)):
        h, w = self.sizes[int(ref)]
        out = _lib.ViewOut()
        res = {}
        if "depth" in want:
            res["depth"] = np.zeros((h, w), np.float32)
            out.depth = fptr(res["depth"])
        if "normal" in want:
            res["normal"] = np.zeros((h, w, 3), np.float32)
            out.normal = fptr(res["normal"])
        if "cost" in want:
            res["cost"] = np.zeros((h, w), np.float32)
            out.cost = fptr(res["cost"])
        if "consistency" in want:
            res["consistency"] = np.zeros((h, w), np.uint32)
            out.consistency = u32ptr(res["consistency"])
        done = ctypes.c_int32()
        self._check(self.lib.vsm_poll(self.h, ctypes.c_uint64(ticket), ctypes.byref(out),
                                      ctypes.byref(done)))
        return res if done.value else None

    # ------------------------------------------------------------------
    def stats(self):
        s = _lib.Stats()
        self._check(self.lib.vsm_get_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        self._check(self.lib.vsm_reset_stats(self.h))

    def set_timing(self, on=True):
        self._check(self.lib.vsm_set_timing(self.h, 1 if on else 0))

    def synchronize(self):
        self._check(self.lib.vsm_synchronize(self.h))


class SolveBatch:
    """Handle of a staged solve (vsm_batch_begin / _step / _end)."""

    def __init__(self, ctx, handle, views):
        self.ctx, self.handle, self.views = ctx, handle, list(views)
        self.remaining = None

    def step(self, n_halfsteps=1):
        """Run up to n more half-steps over every view; returns how many are left."""
        rem = ctypes.c_int32()
        self.ctx._check(self.ctx.lib.vsm_batch_step(self.ctx.h, self.handle, int(n_halfsteps),
                                                    ctypes.byref(rem)))
        self.remaining = rem.value
        return rem.value

    def end(self, want=("depth", "normal", "cost", "consistency")):
        outs, results = self.ctx._outputs(self.views, want)
        h, self.handle = self.handle, None
        self.ctx._check(self.ctx.lib.vsm_batch_end(self.ctx.h, h, outs, 0))
        return results

    def discard(self):
        if self.handle is not None:
            h, self.handle = self.handle, None
            self.ctx._check(self.ctx.lib.vsm_batch_end(self.ctx.h, h, None, 1))
