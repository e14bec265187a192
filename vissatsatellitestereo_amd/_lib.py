"""ctypes binding of libvissat_mvs.so (C ABI in include/vissat_mvs.h).

The library is built in-tree by ``vissatsatellitestereo_amd.build``.  There is
no CPU fallback: if the library is missing or no HIP device is present, every
entry point raises.
"""
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvissat_mvs.so")

VSM_OK = 0
ERRORS = {-22: "EINVAL", -12: "ENOMEM", -5: "EDEVICE", -2: "ENOENT", -11: "EAGAIN"}
MAX_SRC = 32
MAX_RADIUS = 8


class VsmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [("window_radius", ctypes.c_int32), ("window_step", ctypes.c_int32),
                ("sigma_spatial", ctypes.c_float), ("sigma_color", ctypes.c_float),
                ("ncc_sigma", ctypes.c_float), ("num_samples", ctypes.c_int32),
                ("num_iterations", ctypes.c_int32), ("min_triangulation_angle", ctypes.c_float),
                ("incident_angle_sigma", ctypes.c_float), ("geom_consistency", ctypes.c_int32),
                ("geom_consistency_regularizer", ctypes.c_float),
                ("geom_consistency_max_cost", ctypes.c_float), ("filter", ctypes.c_int32),
                ("filter_min_ncc", ctypes.c_float),
                ("filter_min_triangulation_angle", ctypes.c_float),
                ("filter_geom_consistency_max_cost", ctypes.c_float),
                ("filter_min_num_consistent", ctypes.c_int32), ("depth_min", ctypes.c_float),
                ("depth_max", ctypes.c_float), ("perturbation", ctypes.c_float),
                ("normal_perturbation", ctypes.c_float), ("max_normal_angle", ctypes.c_float),
                ("perturb_decay", ctypes.c_float), ("n_far", ctypes.c_int32),
                ("seed", ctypes.c_uint64), ("draw_group", ctypes.c_int32),
                ("tile_width", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ViewReq(ctypes.Structure):
    _fields_ = [("ref_id", ctypes.c_int32), ("n_src", ctypes.c_int32),
                ("src_ids", ctypes.POINTER(ctypes.c_int32)), ("depth_min", ctypes.c_float),
                ("depth_max", ctypes.c_float), ("init_depth", ctypes.POINTER(ctypes.c_float)),
                ("init_normal", ctypes.POINTER(ctypes.c_float))]


class ViewOut(ctypes.Structure):
    _fields_ = [("depth", ctypes.POINTER(ctypes.c_float)), ("normal", ctypes.POINTER(ctypes.c_float)),
                ("cost", ctypes.POINTER(ctypes.c_float)),
                ("consistency", ctypes.POINTER(ctypes.c_uint32))]


class DsmGrid(ctypes.Structure):
    _fields_ = [("ul_e", ctypes.c_double), ("ul_n", ctypes.c_double), ("e_res", ctypes.c_double),
                ("n_res", ctypes.c_double), ("e_size", ctypes.c_int32), ("n_size", ctypes.c_int32)]


class Fuse3dParams(ctypes.Structure):
    _fields_ = [("min_num_pixels", ctypes.c_int32), ("max_reproj_error", ctypes.c_float),
                ("max_depth_error", ctypes.c_float), ("max_normal_error_deg", ctypes.c_float)]


class GeoParams(ctypes.Structure):
    _fields_ = [("lat0_deg", ctypes.c_double), ("lon0_deg", ctypes.c_double), ("alt0", ctypes.c_double),
                ("zone_number", ctypes.c_int32), ("south", ctypes.c_int32)]


class GeoFrame(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_double * 3), ("rot", ctypes.c_double * 9),
                ("cm_cos", ctypes.c_double), ("cm_sin", ctypes.c_double),
                ("a", ctypes.c_double), ("e", ctypes.c_double), ("e2", ctypes.c_double),
                ("e2a", ctypes.c_double), ("ep2b", ctypes.c_double), ("one_minus_f", ctypes.c_double),
                ("k0A", ctypes.c_double), ("alpha", ctypes.c_double * 6),
                ("false_easting", ctypes.c_double), ("false_northing", ctypes.c_double),
                ("max_q", ctypes.c_double)]

    def as_array(self):
        """The 30 constants in declaration order (FP64)."""
        import numpy as np
        return np.frombuffer(bytes(self), dtype=np.float64).copy()


class Stats(ctypes.Structure):
    _fields_ = [("halfstep_ms", ctypes.c_double), ("halfstep_launches", ctypes.c_int64),
                ("ncc_evals", ctypes.c_int64), ("geo_evals", ctypes.c_int64),
                ("taps", ctypes.c_int64), ("fuse_ms", ctypes.c_double),
                ("fuse_launches", ctypes.c_int64), ("fuse_pixels", ctypes.c_int64),
                ("hbm_bytes_in_use", ctypes.c_uint64), ("hbm_bytes_peak", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# (name, restype, argtypes) of every exported symbol; tests check each exists.
P = ctypes.POINTER
_c = ctypes
SIGNATURES = [
    ("vsm_default_params", None, [P(Params)]),
    ("vsm_create", _c.c_int, [_c.c_int, _c.c_uint64, P(_c.c_void_p)]),
    ("vsm_destroy", _c.c_int, [_c.c_void_p]),
    ("vsm_last_error", _c.c_char_p, [_c.c_void_p]),
    ("vsm_set_params", _c.c_int, [_c.c_void_p, P(Params)]),
    ("vsm_get_params", _c.c_int, [_c.c_void_p, P(Params)]),
    ("vsm_add_image", _c.c_int, [_c.c_void_p, _c.c_int32, P(_c.c_uint8), _c.c_int32, _c.c_int32,
                                 _c.c_int32, P(_c.c_double)]),
    ("vsm_remove_image", _c.c_int, [_c.c_void_p, _c.c_int32]),
    ("vsm_set_depth_map", _c.c_int, [_c.c_void_p, _c.c_int32, _c.c_int32, P(_c.c_float), _c.c_int32,
                                     _c.c_int32]),
    ("vsm_get_depth_map", _c.c_int, [_c.c_void_p, _c.c_int32, _c.c_int32, P(_c.c_float), _c.c_int32,
                                     _c.c_int32]),
    ("vsm_solve", _c.c_int, [_c.c_void_p, P(ViewReq), _c.c_int32, _c.c_int32, P(ViewOut), _c.c_int32]),
    ("vsm_batch_begin", _c.c_int, [_c.c_void_p, P(ViewReq), _c.c_int32, _c.c_int32, _c.c_int32,
                                   P(_c.c_int32)]),
    ("vsm_batch_step", _c.c_int, [_c.c_void_p, _c.c_int32, _c.c_int32, P(_c.c_int32)]),
    ("vsm_batch_end", _c.c_int, [_c.c_void_p, _c.c_int32, P(ViewOut), _c.c_int32]),
    ("vsm_set_hbm_budget", _c.c_int, [_c.c_void_p, _c.c_uint64]),
    ("vsm_eval_cost_staging", _c.c_int, [_c.c_void_p, P(ViewReq), _c.c_int32, P(_c.c_float),
                                         P(_c.c_float), _c.c_int32, _c.c_int32, P(_c.c_float),
                                         P(_c.c_float), P(_c.c_float)]),
    ("vsm_proj_to_grid_noalloc", _c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_int64, P(DsmGrid),
                                            _c.c_void_p, _c.c_void_p]),
    ("vsm_submit", _c.c_int, [_c.c_void_p, P(ViewReq), P(_c.c_uint64)]),
    ("vsm_submit_pass", _c.c_int, [_c.c_void_p, P(ViewReq), _c.c_int32, _c.c_int32, P(_c.c_uint64)]),
    ("vsm_serve_step", _c.c_int, [_c.c_void_p, P(_c.c_int32), P(_c.c_int32)]),
    ("vsm_poll", _c.c_int, [_c.c_void_p, _c.c_uint64, P(ViewOut), P(_c.c_int32)]),
    ("vsm_eval_cost", _c.c_int, [_c.c_void_p, P(ViewReq), _c.c_int32, P(_c.c_float), P(_c.c_float),
                                 P(_c.c_float)]),
    ("vsm_debug_halfstep", _c.c_int, [_c.c_void_p, P(ViewReq), _c.c_int32, _c.c_int32, _c.c_int32,
                                      P(_c.c_float), P(_c.c_float), P(_c.c_float), P(_c.c_float)]),
    ("vsm_debug_init", _c.c_int, [_c.c_void_p, P(ViewReq), _c.c_int32, P(_c.c_float), P(_c.c_float),
                                  P(_c.c_float), P(_c.c_float)]),
    ("vsm_debug_filter", _c.c_int, [_c.c_void_p, P(ViewReq), _c.c_int32, P(_c.c_float), P(_c.c_float),
                                    P(_c.c_uint32)]),
    ("vsm_fuse_dsm", _c.c_int, [_c.c_void_p, P(DsmGrid), P(_c.c_int32), _c.c_int32, _c.c_int32,
                                _c.c_int32, P(_c.c_double), P(_c.c_float), P(_c.c_float)]),
    ("vsm_proj_to_grid", _c.c_int, [_c.c_void_p, P(_c.c_double), _c.c_int64, P(DsmGrid),
                                    P(_c.c_double)]),
    ("vsm_median_blur3", _c.c_int, [_c.c_void_p, P(_c.c_float), _c.c_int32, _c.c_int32,
                                    P(_c.c_float)]),
    ("vsm_set_normal_map", _c.c_int, [_c.c_void_p, _c.c_int32, _c.c_int32, P(_c.c_float), _c.c_int32,
                                      _c.c_int32]),
    ("vsm_default_fuse3d_params", None, [P(Fuse3dParams)]),
    ("vsm_fuse_points", _c.c_int, [_c.c_void_p, P(_c.c_int32), _c.c_int32, _c.c_int32, P(Fuse3dParams),
                                   P(_c.c_double), P(_c.c_double), _c.c_int64, P(_c.c_double),
                                   P(_c.c_float), P(_c.c_uint8), P(_c.c_int64)]),
    ("vsm_geo_frame_init", _c.c_int, [P(GeoParams), P(GeoFrame)]),
    ("vsm_set_geo_frame", _c.c_int, [_c.c_void_p, P(GeoFrame)]),
    ("vsm_enu_to_utm", _c.c_int, [_c.c_void_p, P(_c.c_double), _c.c_int64, P(_c.c_double)]),
    ("vsm_get_stats", _c.c_int, [_c.c_void_p, P(Stats)]),
    ("vsm_reset_stats", _c.c_int, [_c.c_void_p]),
    ("vsm_set_timing", _c.c_int, [_c.c_void_p, _c.c_int32]),
    ("vsm_synchronize", _c.c_int, [_c.c_void_p]),
]

def geo_frame(lat0_deg, lon0_deg, alt0, zone_number, south):
    """vsm_geo_frame_init: the FP64 constants of the ENU -> UTM chain."""
    p = GeoParams(float(lat0_deg), float(lon0_deg), float(alt0), int(zone_number), int(south))
    f = GeoFrame()
    rc = load().vsm_geo_frame_init(ctypes.byref(p), ctypes.byref(f))
    if rc != 0:
        raise ValueError(f"vsm_geo_frame_init rejected {lat0_deg, lon0_deg, alt0, zone_number, south}")
    return f


_LIB = None
_LOCK = threading.Lock()


def load(path=LIB_PATH):
    """Load the library (no device access).  Raises if it is not built."""
    global _LIB
    with _LOCK:
        if _LIB is None:
            if not os.path.exists(path):
                raise RuntimeError(f"{path} is missing: build it with "
                                   "`python -m vissatsatellitestereo_amd.build`")
            lib = ctypes.CDLL(path)
            for name, res, args in SIGNATURES:
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _LIB = lib
    return _LIB


def default_params(**kw):
    p = Params()
    load().vsm_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if a is not None else None


def u32ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) if a is not None else None


def make_req(ref_id, src_ids, depth_range, init_depth=None, init_normal=None):
    """ViewReq plus the arrays it points to (keep the tuple alive)."""
    ids = np.ascontiguousarray(src_ids, dtype=np.int32)
    keep = [ids]
    r = ViewReq()
    r.ref_id = int(ref_id)
    r.n_src = len(ids)
    r.src_ids = ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    r.depth_min, r.depth_max = float(depth_range[0]), float(depth_range[1])
    if init_depth is not None:
        d = np.ascontiguousarray(init_depth, dtype=np.float32)
        n = np.ascontiguousarray(init_normal, dtype=np.float32)
        keep += [d, n]
        r.init_depth = fptr(d)
        r.init_normal = fptr(n)
    return r, keep
