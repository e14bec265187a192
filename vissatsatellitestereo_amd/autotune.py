"""Autotuner for the batched form's window schedule.

Knob: the tap stride inside the matching window (``window_step``; the
candidates are 1 and 2 by default, stride s visits (floor(2r/s)+1)^2 taps
instead of (2r+1)^2).  Wider strides can be offered as candidates (``steps``);
the accuracy proxy below decides whether they are accepted.

Accuracy proxy (no ground truth needed): multi-view geometric consistency.
Each candidate solves a calibration batch (up to ``calib_views`` reference
views, sources restricted to that batch) through the photometric pass, the
geometric pass and the filter; the fraction of pixels the filter keeps is the
share of depths that at least two sources confirm photometrically and by
forward-backward reprojection.  A candidate is accepted when its kept
fraction is within ``max_drop`` of stride 1's; the fastest accepted one (wall
time of both passes) wins.  Results are cached in a JSON file keyed by
(height, width, sources, radius, iterations) size classes and the candidate set.

Second knob: the half-step kernel's tile shape (``tile_width`` 16, 32 or 64:
tiles of 16 x 32, 32 x 16 or 64 x 8 pixels; a wave covers 8, 4 or 2 tile
rows, which changes the footprint its gathers touch in each source image).
The shape is a schedule only (results agree to FP32 rounding), so it is
chosen by time alone: with the chosen window step, a few views of the batch
with their full source lists run eight photometric half-steps per candidate,
the last four timed (past the first iterations' random hypotheses); the
fastest wins.  64-wide tiles are offered only for views with at most 16
sources.  With grouped source draws (draw_group = 1) the draw groups follow
the tile shape, so the candidates are timed with their own grouping.
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


TILE_WIDTHS = (16, 32, 64)


def key_of(shape, n_src, radius, iters, steps=(1, 2), widths=TILE_WIDTHS):
    cand = "".join(str(int(s)) for s in steps)
    tw = "-".join(str(int(w)) for w in widths)
    return f"h{size_class(shape[0])}_w{size_class(shape[1])}_s{n_src}_r{radius}_i{iters}_c{cand}_t{tw}"


def tune_tiles(ctx, views, step, widths=TILE_WIDTHS, n_views=4, halfsteps=4):
    """Seconds per candidate tile width: photometric half-steps
    halfsteps .. 2 halfsteps - 1 of the first n_views views (full source
    lists), after the first `halfsteps` untimed."""
    calib = views[:n_views]
    n_src = max(len(v["srcs"]) for v in calib)
    saved = ctx.params
    times = {}
    try:
        for tw in widths:
            if tw == 64 and n_src > 16:
                continue
            ctx.set_params(window_step=step, tile_width=tw, geom_consistency=0, filter=0,
                           num_iterations=max(int(saved.num_iterations), halfsteps))
            b = ctx.batch(calib, 0)
            b.step(halfsteps)                                     # random start, warm caches
            ctx.synchronize()
            t = time.perf_counter()
            b.step(halfsteps)
            ctx.synchronize()
            times[tw] = time.perf_counter() - t
            b.discard()
    finally:
        ctx.set_params(window_step=saved.window_step, tile_width=saved.tile_width,
                       geom_consistency=saved.geom_consistency, filter=saved.filter,
                       num_iterations=saved.num_iterations)
    return times


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


def tune(ctx, views, iters, radius, steps=(1, 2), max_drop=0.02, calib_views=8,
         cache_path=CACHE, use_cache=True, widths=TILE_WIDTHS):
    """Pick window_step and tile_width for `views` on `ctx`; returns
    (step, report) with report["tile_width"] the chosen width.

    Note: the calibration solves replace the context's stored photometric
    and geometric maps of the calibration views."""
    shape = ctx.sizes[int(views[0]["ref"])]
    n_src = max(len(v["srcs"]) for v in views)
    steps = tuple(int(s) for s in steps if 1 <= int(s) <= radius)
    if not steps or steps[0] != 1:
        raise ValueError("window-step candidates must start with 1 (the accuracy reference)")
    widths = tuple(int(w) for w in widths)
    key = key_of(shape, n_src, radius, iters, steps, widths)
    cache = load_cache(cache_path) if use_cache else {}
    if key in cache:
        return cache[key]["window_step"], dict(cache[key], cached=True)
    ids = {int(v["ref"]) for v in views[:calib_views]}
    calib = []
    for v in views[:calib_views]:
        srcs = [s for s in v["srcs"] if int(s) in ids]
        if srcs:
            calib.append(dict(v, srcs=srcs))
    saved = ctx.params
    kept, times = {}, {}
    try:
        for st in steps:
            ctx.set_params(window_radius=radius, num_iterations=iters, window_step=st,
                           geom_consistency=0, filter=0)
            ctx.solve(calib[:1], 0, want=())                     # warm-up
            ctx.synchronize()
            t = time.perf_counter()
            ctx.solve(calib, 0, want=())
            ctx.set_params(geom_consistency=1, filter=1)
            res = ctx.solve(calib, 1, want=("depth",))
            ctx.synchronize()
            times[st] = time.perf_counter() - t
            kept[st] = float(np.mean([np.mean(r["depth"] > 0) for r in res]))
    finally:
        ctx.set_params(window_radius=saved.window_radius, window_step=saved.window_step,
                       num_iterations=saved.num_iterations,
                       geom_consistency=saved.geom_consistency, filter=saved.filter)
    ok = [st for st in steps if kept[st] >= kept[steps[0]] - max_drop]
    best = min(ok, key=lambda st: times[st])
    tile_times = tune_tiles(ctx, views, best, widths) if len(widths) > 1 else {}
    tile = min(tile_times, key=lambda w: tile_times[w]) if tile_times else int(ctx.params.tile_width)
    rep = dict(window_step=best, tile_width=tile, key=key, seconds={str(k): v for k, v in times.items()},
               kept_fraction={str(k): v for k, v in kept.items()}, max_drop=max_drop,
               tile_seconds={str(k): v for k, v in tile_times.items()},
               calib_views=len(calib), cached=False)
    if use_cache:
        cache[key] = rep
        save_cache(cache, cache_path)
    return best, rep
