"""Drop-in for the reference's colmap_mvs_commands.py.

``run_photometric_mvs(mvs_dir, window_radius, depth_range=None)`` and
``run_consistency_check(mvs_dir, window_radius, depth_range=None)`` keep the
reference signatures (colmap_mvs_commands.py:39, 58) and flag values
(colmap_mvs_commands.py:40-48, 60-75), and the COLMAP dense workspace they
read and write:

* in:  images/, sparse/ (via last_rows.txt -> P4), depth_ranges.txt,
       stereo/patch-match.cfg;
* out: stereo/depth_maps/<img>.{photometric,geometric}.bin,
       stereo/normal_maps/<img>.{photometric,geometric}.bin (COLMAP array
       format, colmap/read_dense.py:36-51), stereo/consistency_graphs/
       <img>.geometric.bin (COLMAP consistency-graph records, see
       colmap/consistency_graph.py) and <img>.geometric.npy (the same as a
       uint32 bitmask of consistent sources per pixel), and inv_proj_mats.txt
       (aggregate_2p5d_util.py:55-61).

Instead of spawning `colmap patch_match_stereo` (lib/run_cmd.py:44) the
views are solved in-process by the HIP library.  Errors raise (the
reference ignores the subprocess exit code, lib/run_cmd.py:44-63).

Extra keyword arguments (defaults = reference behaviour):
  num_iterations  12 photometric / 1 geometric (colmap_mvs_commands.py:47, 75)
  batch_size      None = all reference views in one batched launch sequence;
                  1 = one view at a time (the plain form)
  use_graph       replay the half-step sequence from a captured hipGraph
  params          further PatchMatchStereo options (vsm_params field names)
"""
import os

import numpy as np

from .colmap import consistency_graph
from .colmap.read_dense import write_array
from . import runtime
from .workspace import select_sources


def _views(ws, depth_range, min_tri):
    srcs = select_sources(ws.mvs_dir, min_triangulation_angle=min_tri)
    dr = ws.depth_ranges()
    views = []
    for ref in ws.names:
        if ref not in srcs or not srcs[ref]:
            continue
        rng = depth_range if depth_range is not None else dr[ref]
        views.append((ref, srcs[ref], rng))
    return views


def _solve(ws, views, pass_, batch_size, use_graph, want, extra):
    ctx = runtime.context()
    out = {}
    step = batch_size or len(views)
    for i in range(0, len(views), step):
        chunk = views[i:i + step]
        req = []
        for ref, srcs, rng in chunk:
            v = dict(ref=ws.ids[ref], srcs=[ws.ids[s] for s in srcs], depth_range=rng)
            if extra:
                v.update(extra(ref))
            req.append(v)
        res = ctx.solve(req, pass_=pass_, want=want, use_graph=use_graph)
        for (ref, _, _), r in zip(chunk, res):
            out[ref] = r
    return out


def run_photometric_mvs(mvs_dir, window_radius, depth_range=None, num_iterations=12,
                        batch_size=None, use_graph=False, params=None):
    ws = runtime.Workspace.get(mvs_dir)
    ctx = runtime.context()
    p = dict(window_radius=int(window_radius), min_triangulation_angle=5.0, filter=0,
             geom_consistency=0, num_samples=15, num_iterations=int(num_iterations))
    p.update(params or {})
    ctx.set_params(**p)
    views = _views(ws, depth_range, p["min_triangulation_angle"])
    needed = sorted({v[0] for v in views} | {s for v in views for s in v[1]})
    ws.upload(ctx, needed)
    ws.write_inv_proj_mats()
    res = _solve(ws, views, 0, batch_size, use_graph, ("depth", "normal"), None)
    for ref, r in res.items():
        write_array(r["depth"], ws.depth_path(ref, "photometric"))
        write_array(r["normal"], ws.normal_path(ref, "photometric"))
        ws.photo_resident.add(ref)
    return res


def run_consistency_check(mvs_dir, window_radius, depth_range=None, num_iterations=1,
                          batch_size=None, use_graph=False, params=None):
    ws = runtime.Workspace.get(mvs_dir)
    ctx = runtime.context()
    p = dict(window_radius=int(window_radius), min_triangulation_angle=5.0, geom_consistency=1,
             geom_consistency_regularizer=0.3, geom_consistency_max_cost=3.0, filter=1,
             filter_min_triangulation_angle=4.99, filter_min_ncc=0.1,
             filter_geom_consistency_max_cost=1.0, filter_min_num_consistent=2, num_samples=15,
             num_iterations=int(num_iterations))
    p.update(params or {})
    ctx.set_params(**p)
    views = _views(ws, depth_range, p["min_triangulation_angle"])
    needed = sorted({v[0] for v in views} | {s for v in views for s in v[1]})
    ws.upload(ctx, needed)
    # use_exist_photom (colmap_mvs_commands.py:64): photometric maps of every
    # image involved, from the device when resident, else from the files.
    for n in needed:
        if n not in ws.photo_resident:
            ctx.set_depth_map(ws.ids[n], ws.read_depth(n, "photometric"), kind=0)

    def init(ref):
        if ref in ws.photo_resident:
            return {}
        return dict(init_depth=ws.read_depth(ref, "photometric"),
                    init_normal=ws.read_normal(ref, "photometric"))

    res = _solve(ws, views, 1, batch_size, use_graph, ("depth", "normal", "consistency"), init)
    cg = os.path.join(ws.mvs_dir, "stereo", "consistency_graphs")
    os.makedirs(cg, exist_ok=True)
    for ref, r in res.items():
        write_array(r["depth"], ws.depth_path(ref, "geometric"))
        write_array(r["normal"], ws.normal_path(ref, "geometric"))
        np.save(os.path.join(cg, f"{ref}.geometric.npy"), r["consistency"])
        srcs = next(v[1] for v in views if v[0] == ref)
        consistency_graph.write(os.path.join(cg, f"{ref}.geometric.bin"), r["consistency"],
                                [ws.names.index(sn) for sn in srcs])
    return res
