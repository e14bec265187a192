"""Autotuner for the batched form's window schedule.

Knob: the tap stride inside the matching window (``window_step`` 1 or 2;
stride 2 visits (floor(2r/2)+1)^2 taps instead of (2r+1)^2).  A candidate is
accepted when the depth maps it produces on a calibration batch agree with
the stride-1 maps: at least ``min_agreement`` of the pixels that stride 1
matched confidently (cost < 0.5) lie within ``depth_tol`` of the stride-1
depth.  Among accepted candidates the fastest by device time wins.  Results
are cached in a JSON file keyed by (height, width, sources, radius,
iterations) size classes.
"""
import json
import os
import time

import numpy as np

CACHE = os.environ.get("VSM_AUTOTUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "vissat_mvs_autotune.json"))


def size_class(n):
    """Power-of-two bucket of an image side."""
    return 1 << int(np.ceil(np.log2(max(1, n))))


def key_of(shape, n_src, radius, iters):
    return f"h{size_class(shape[0])}_w{size_class(shape[1])}_s{n_src}_r{radius}_i{iters}"


def load_cache(path=CACHE):
    try:
        with open(path) as fp:
            return json.load(fp)
    except (OSError, ValueError):
        return {}


def save_cache(cache, path=CACHE):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as fp:
        json.dump(cache, fp, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune(ctx, views, iters, radius, steps=(1, 2), min_agreement=0.9, depth_tol=0.5,
         calib_views=2, cache_path=CACHE, use_cache=True):
    """Pick window_step for `views` on `ctx`; returns (step, report)."""
    shape = ctx.sizes[int(views[0]["ref"])]
    n_src = max(len(v["srcs"]) for v in views)
    key = key_of(shape, n_src, radius, iters)
    cache = load_cache(cache_path) if use_cache else {}
    if key in cache:
        return cache[key]["window_step"], dict(cache[key], cached=True)
    calib = views[:calib_views]
    saved = ctx.params
    ctx.set_params(window_radius=radius, num_iterations=iters, geom_consistency=0, filter=0)
    results, times = {}, {}
    try:
        for st in steps:
            ctx.set_params(window_step=st)
            ctx.solve(calib, 0, want=("depth",))            # warm-up (code objects, allocations)
            ctx.synchronize()
            t = time.perf_counter()
            results[st] = ctx.solve(calib, 0, want=("depth", "cost"))
            ctx.synchronize()
            times[st] = time.perf_counter() - t
    finally:
        ctx.set_params(window_step=saved.window_step, num_iterations=saved.num_iterations,
                       geom_consistency=saved.geom_consistency, filter=saved.filter)
    ref = results[steps[0]]
    agree = {}
    for st in steps:
        fr = []
        for a, b in zip(ref, results[st]):
            conf = a["cost"] < 0.5
            fr.append(float(np.mean(np.abs(a["depth"] - b["depth"])[conf] < depth_tol)) if conf.any() else 1.0)
        agree[st] = float(np.mean(fr))
    ok = [st for st in steps if agree[st] >= min_agreement]
    best = min(ok, key=lambda st: times[st]) if ok else steps[0]
    rep = dict(window_step=best, key=key, seconds={str(k): v for k, v in times.items()},
               agreement={str(k): v for k, v in agree.items()}, min_agreement=min_agreement,
               depth_tol=depth_tol, cached=False)
    if use_cache:
        cache[key] = rep
        save_cache(cache, cache_path)
    return best, rep
