"""MVS throughput benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1|c2|c3|c4|c5]
                    [--batched-step auto|1|2] [--tune-steps 1,2] [--no-cpu-baseline]

Configurations (synthetic sites, seeded; images resident in HBM before
timing; one JSON line on stdout):

c2 (default)  10 views of 1024x1024, every view a reference with 8 sources;
              photometric pass (5 iterations), geometric pass + filter
              (5 iterations), 0.5 m DSM fusion.  One step solves and fuses
              all 10 views.  Two forms on the same inputs:
                plain    one reference view per solve call, synchronised,
                         window step 1;
                batched  all views in one solve per pass, the half-step
                         sequence replayed from a captured hipGraph, window
                         step chosen by the autotuner.
              `value` = batched views/s; plain views/s, per-view latency
              percentiles and the DSM error of both forms vs the synthetic
              ground truth are reported beside it.
c3            32 reference views of ragged crop sizes (512..2048 per side),
              6..12 sources each, same passes; plain vs batched.
c4            serving: Poisson arrivals (default 20 req/s) of per-view depth
              requests from 4 sites (crops 1024..2048 per side, 8..11
              sources); continuous batching with HBM-footprint admission;
              p50/p99 latency and views/s, against a one-request-at-a-time
              queue on the same arrival schedule; `capacity` repeats both
              forms with every request queued at t = 0.
c5            40 views of 4096x4096, 20 sources, both passes, 0.3 m DSM.
c1            2 views of 256x256, 1 source, photometric only (the case the
              CPU oracle runs in full).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = "MVS views/s + p50/p99 per-view latency, batched vs plain, 1 MI355X"
FP32_PEAK_TFLOPS = 157.3      # MI355X vector FP32 (published)
HBM_PEAK_GBPS = 8000.0        # MI355X HBM3E (published)
FLOPS_PER_TAP = 30            # homography 12, divide 3, bilinear 8, accumulate 7


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# sites
# ---------------------------------------------------------------------------
def _device():
    import torch
    return "cuda" if torch.cuda.is_available() else "cpu"


def build_site(spec, work, max_src, src_counts=None, dsm_resolution=0.5):
    """Render + dense workspace; src_counts: optional per-view source counts
    written as explicit patch-match.cfg lists."""
    from vissatsatellitestereo_amd import synth, workspace
    if not os.path.exists(os.path.join(work, "colmap", "mvs", "depth_ranges.txt")):
        t = time.time()
        synth.build_site(spec, work, device=_device(), save_gt_views=False, log=log,
                         dsm_resolution=dsm_resolution)
        mvs = workspace.prepare_dense_workspace(work, max_src=max_src)
        if src_counts is not None:
            sel = workspace.select_sources(mvs)
            names = sorted(sel)
            with open(os.path.join(mvs, "stereo", "patch-match.cfg"), "w") as fp:
                for n, k in zip(names, src_counts):
                    fp.write(n + "\n" + ", ".join(sel[n][:k]) + "\n")
        log(f"site {work} built in {time.time() - t:.1f}s")
    return work


def load_site(work, id_base=0):
    from vissatsatellitestereo_amd import reparam_depth as rd
    from vissatsatellitestereo_amd.workspace import select_sources
    from PIL import Image
    mvs = os.path.join(work, "colmap", "mvs")
    P4 = rd.projection_matrices(mvs)
    names = sorted(P4)
    imgs = {n: np.array(Image.open(os.path.join(mvs, "images", n))) for n in names}
    dr = rd.read_name_values(os.path.join(mvs, "depth_ranges.txt"))
    srcs = select_sources(mvs)
    ids = {n: id_base + i + 1 for i, n in enumerate(names)}
    views = [dict(ref=ids[n], srcs=[ids[s] for s in srcs[n]], depth_range=dr[n]) for n in names]
    return dict(work=work, names=names, imgs=imgs, P4=P4, views=views, ids=ids)


def dsm_error(dsm, gt):
    ok = ~np.isnan(dsm) & ~np.isnan(gt)
    if not ok.any():
        return None, 0.0
    return float(np.median(np.abs(dsm[ok] - gt[ok]))), float(ok.mean())


# ---------------------------------------------------------------------------
# plain vs batched
# ---------------------------------------------------------------------------
PASSES_FULL = ((0, dict(geom_consistency=0, filter=0)), (1, dict(geom_consistency=1, filter=1)))
PASSES_PHOTO = ((0, dict(geom_consistency=0, filter=0)),)


def run_form(ctx, views, grid, form, iters, use_graph, photometric_only=False):
    """One step: photometric (+ geometric with filter) for every view, then
    fusion.  Returns (wall seconds, per-view latencies, fused DSM)."""
    lat = {}
    t0 = time.perf_counter()
    for pass_, extra in (PASSES_PHOTO if photometric_only else PASSES_FULL):
        ctx.set_params(num_iterations=iters, **extra)
        if form == "plain":
            for v in views:
                ts = time.perf_counter()
                ctx.solve([v], pass_, want=())
                lat[v["ref"]] = lat.get(v["ref"], 0.0) + time.perf_counter() - ts
        else:
            ctx.solve(views, pass_, want=(), use_graph=use_graph)
    # two views (c1) cannot pass the reference's <= 2 rejection: keep every cell there
    dsm = ctx.fuse_dsm(grid, [v["ref"] for v in views], depth_kind=0 if photometric_only else 1,
                       min_measurements=1 if photometric_only else 3)
    wall = time.perf_counter() - t0
    if form != "plain":
        lat = {v["ref"]: wall for v in views}
    return wall, list(lat.values()), dsm


def measure_forms(ctx, views, grid, gt, a, batched_step):
    res = {}
    for form in [f for f in ("plain", "batched") if f in a.forms.split(",")]:
        ctx.set_params(window_step=1 if form == "plain" else batched_step)
        use_graph = form == "batched"
        po = getattr(a, "photometric_only", False)
        for _ in range(a.warmup):
            run_form(ctx, views, grid, form, a.iters, use_graph, po)
        log(f"{form}: warm")
        ctx.reset_stats()
        ctx.set_timing(form == "batched")
        walls, lats = [], []
        ctx.synchronize()
        for k in range(a.steps):
            w, lat, dsm = run_form(ctx, views, grid, form, a.iters, use_graph, po)
            walls.append(w)
            lats += lat
            log(f"{form} step {k}: {w * 1e3:.0f} ms")
        ctx.synchronize()
        st = ctx.stats()
        ctx.set_timing(False)
        err, cov = dsm_error(dsm, gt)
        res[form] = dict(ms_per_step=1e3 * float(np.mean(walls)),
                         views_per_s=len(views) * a.steps / float(np.sum(walls)),
                         p50_ms=1e3 * float(np.percentile(lats, 50)),
                         p99_ms=1e3 * float(np.percentile(lats, 99)),
                         dsm_err=err, dsm_cov=cov, stats=st)
        log(form, {k: v for k, v in res[form].items() if k != "stats"})
    return res


def roofline(st, cfg="c2"):
    if not (st["halfstep_ms"] > 0 and st["halfstep_launches"] > 0):
        return None
    avg_ms = st["halfstep_ms"] / st["halfstep_launches"]
    taps = st["taps"] / st["halfstep_launches"]
    ach = FLOPS_PER_TAP * taps / (avg_ms * 1e-3) / 1e12
    traffic, source, profiled = None, None, None
    # HBM traffic per launch comes from a committed counter profile of the
    # same configuration (none for the others)
    tfile = os.path.join(ROOT, "profiles", f"k_halfstep_{cfg}_counters.json")
    if os.path.exists(tfile):
        t = json.load(open(tfile))
        traffic = round(t["fetch_bytes"] + t["write_bytes"])
        source = (f"profiles/k_halfstep_{cfg}_counters.md: FETCH_SIZE + WRITE_SIZE per dispatch, "
                  "calibrated on an elementwise kernel of known bytes (batched, window step 2)"
                  + (f"; {t['calibration']}" if "calibration" in t else ""))
        # the profile's traffic over this run's live launch time, against the
        # 8 TB/s HBM peak; VALU issue and lane figures as the profile read them
        gbps = traffic / (avg_ms * 1e-3) / 1e9
        profiled = {"hbm_gbps": round(gbps, 1), "hbm_frac": round(gbps / HBM_PEAK_GBPS, 4),
                    "valu_busy_pct": round(t["valubusy"], 1) if "valubusy" in t else None,
                    "valu_lanes_active_pct": round(t["valuutil"], 1) if "valuutil" in t else None}
    return {"bound": "valu", "kernel": "k_halfstep", "achieved": round(ach, 3),
            "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(ach / FP32_PEAK_TFLOPS, 4),
            "traffic": traffic, "traffic_unit": "bytes per launch", "traffic_source": source,
            "profiled": profiled,
            "bound_note": "vector FP32 (VALU): a gather/stencil kernel, neither MFMA nor HBM "
                          "work; peak = MI355X FP32 157.3 TFLOP/s",
            "avg_launch_ms": round(avg_ms, 4), "taps_per_launch": int(taps),
            "flops_per_tap": FLOPS_PER_TAP}


def cpu_baseline(site, window_radius, iters):
    """Oracle (FP64 C, OpenMP) on a bounded sample: a 512x512 crop of the
    first reference view with all its sources, 1 photometric iteration,
    extrapolated to a full view (area x iterations of both passes)."""
    import pm_oracle as po
    v = site["views"][0]
    inv_ids = {i: n for n, i in site["ids"].items()}
    ref = inv_ids[v["ref"]]
    srcs = [inv_ids[s] for s in v["srcs"]]
    imgs, P4 = site["imgs"], site["P4"]
    H, W = imgs[ref].shape
    ch, cw = min(512, H), min(512, W)
    y0, x0 = (H - ch) // 2, (W - cw) // 2
    crop = np.ascontiguousarray(imgs[ref][y0:y0 + ch, x0:x0 + cw])
    T = np.array([[1, 0, x0, 0], [0, 1, y0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], dtype=np.float64)
    P4c = np.linalg.inv(T) @ P4[ref]          # pixel (c, r) of the crop is (c + x0, r + y0)
    ov = po.OracleView(crop, [imgs[s] for s in srcs], P4c, [P4[s] for s in srcs], v["depth_range"],
                       view_id=v["ref"], params=dict(window_radius=window_radius))
    t = time.perf_counter()
    ov.solve(0, 1)
    dt = time.perf_counter() - t
    scale = (H * W) / (ch * cw)
    per_view = dt * scale * (2 * iters)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(1.0 / per_view, 6), "unit": "views/s", "cores": cores, "kind": "port",
            "sample": f"{ch}x{cw} crop of one {W}x{H} reference view, {len(srcs)} sources, "
                      f"1 photometric iteration of the FP64 oracle in {dt:.1f}s; per-view time = "
                      f"sample x {scale:.0f} (area) x {2 * iters} (iterations of both passes)"}


def cpu_baseline_full(site, window_radius, iters):
    """Oracle (FP64 C, OpenMP) on the whole c1 workload: every view, all
    iterations, no extrapolation."""
    import pm_oracle as po
    inv_ids = {i: n for n, i in site["ids"].items()}
    imgs, P4 = site["imgs"], site["P4"]
    t = time.perf_counter()
    for v in site["views"]:
        ref = inv_ids[v["ref"]]
        srcs = [inv_ids[s] for s in v["srcs"]]
        ov = po.OracleView(imgs[ref], [imgs[s] for s in srcs], P4[ref], [P4[s] for s in srcs],
                           v["depth_range"], view_id=v["ref"], params=dict(window_radius=window_radius))
        ov.solve(0, iters)
    dt = time.perf_counter() - t
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n = len(site["views"])
    return {"value": round(n / dt, 4), "unit": "views/s", "cores": cores, "kind": "port",
            "sample": f"the whole workload: {n} views, {iters} photometric iterations, in {dt:.2f}s"}


def cpu_baseline_fusion(ctx, site, grid, max_views=2, band_rows=256):
    """numpy restatement of aggregate_2p5d (oracle/fusion_oracle.py), one
    process, on the geometric depth maps of the last batched step: per-view
    DSMs timed on up to max_views views and scaled to all of them, the
    cross-view fusion timed on a band of band_rows grid rows and scaled to
    the whole grid (the numpy median of big stacks runs for minutes)."""
    import fusion_oracle as fo
    ids = [v["ref"] for v in site["views"]]
    k = min(max_views, len(ids))
    t = time.perf_counter()
    per_view = []
    for i, n in zip(ids[:k], site["names"][:k]):
        per_view.append(fo.view_dsm(ctx.get_depth_map(i, kind=1), np.linalg.inv(site["P4"][n]), grid))
        log(f"cpu fusion baseline: per-view DSM {n} done")
    t_view = (time.perf_counter() - t) / k
    rows = min(band_rows, grid[5])
    band = np.stack([pv[:rows] for pv in per_view] * (len(ids) // k + 1))[:len(ids)]
    t = time.perf_counter()
    fo.fuse(band, 3)
    t_fuse = (time.perf_counter() - t) * grid[5] / rows
    total = t_view * len(ids) + t_fuse
    return {"value": round(len(ids) / total, 5), "unit": "views/s (fusion only)", "cores": 1,
            "kind": "port",
            "sample": f"per-view DSM of {k} of {len(ids)} geometric depth maps ({t_view:.2f}s each) "
                      f"x {len(ids)}, plus the {len(ids)}-view fusion on {rows} of {grid[5]} grid rows "
                      f"scaled by rows ({t_fuse:.2f}s)"}


def make_ctx(sites, a):
    from vissatsatellitestereo_amd.mvs import MvsContext
    ctx = MvsContext(device=0, hbm_budget_bytes=int(a.hbm_budget_gb * 1e9),
                     window_radius=a.window_radius, window_step=1, num_samples=15)
    for s in sites:
        for n in s["names"]:
            ctx.add_image(s["ids"][n], s["imgs"][n], s["P4"][n])
    ctx.synchronize()
    return ctx


def choose_step(ctx, views, a):
    if a.batched_step != "auto":
        return int(a.batched_step), {"window_step": int(a.batched_step), "source": "command line"}
    from vissatsatellitestereo_amd import autotune
    steps = tuple(int(x) for x in a.tune_steps.split(","))
    st, rep = autotune.tune(ctx, views, iters=a.iters, radius=a.window_radius, steps=steps,
                            use_cache=not a.no_tune_cache)
    log("autotune", rep)
    return st, rep


def bench_forms(a, cfg):
    from vissatsatellitestereo_amd import synth, runtime
    res_m = 0.5
    a.photometric_only = cfg == "c1"
    if cfg == "c1":
        spec, max_src, counts = synth.spec_c1(), 1, None
        desc = "2 views x 256x256 uint8, 1 source each, photometric only"
    elif cfg == "c2":
        spec, max_src, counts = synth.spec_c2(), 8, None
        desc = "10 reference views x 1024x1024 uint8, 8 sources"
    elif cfg == "c5":
        spec, max_src, counts = synth.spec_c5(), 20, None
        desc = "40 reference views x 4096x4096 uint8, 20 sources, 500 buildings"
        res_m = 0.3
    else:
        spec = synth.spec_ragged(32)
        rng = np.random.default_rng(7)
        counts = [int(k) for k in rng.integers(6, 13, spec.n_views)]
        max_src = 12
        desc = "32 reference views, ragged crops 512..2048 per side, 6..12 sources"
    work = build_site(spec, os.path.join(a.work, cfg), max_src, counts, res_m)
    if a.build_only:
        return None
    site = load_site(work)
    gt = np.load(os.path.join(work, "synthetic_gt", "gt_dsm.npy"))
    grid = runtime.dsm_grid(runtime.read_aoi(work), res_m, res_m)
    copy_calibration(a)
    ctx = make_ctx([site], a)
    views = site["views"]
    step, tune_rep = choose_step(ctx, views, a)
    res = measure_forms(ctx, views, grid, gt, a, step)
    b = res.get("batched") or res["plain"]
    p = res.get("plain") or b
    if a.no_cpu_baseline:
        cpu = cpu_fuse = None
    elif cfg == "c1":
        cpu = cpu_baseline_full(site, a.window_radius, a.iters)
        cpu_fuse = None
    else:
        log("cpu baseline: PatchMatch oracle on a crop")
        cpu = cpu_baseline(site, a.window_radius, a.iters)
        log("cpu baseline: fusion")
        cpu_fuse = cpu_baseline_fusion(ctx, site, grid)
    out = {
        "metric": METRIC, "value": round(b["views_per_s"], 4), "unit": "views/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(b["ms_per_step"], 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{cfg}: {desc}, photometric+geometric {a.iters}+{a.iters} "
                               f"iterations, filter, r={a.window_radius}, {res_m} m DSM fusion",
                   "forms": {"plain": "one view per solve call, window step 1",
                             "batched": f"all views per solve call, persistent kernel, hipGraph "
                                        f"replay, window step {step} (autotuner)"},
                   "autotune": tune_rep},
        "plain_views_per_s": round(p["views_per_s"], 4),
        "batched_over_plain": round(b["views_per_s"] / p["views_per_s"], 3),
        "latency_ms": {"plain_p50": round(p["p50_ms"], 1), "plain_p99": round(p["p99_ms"], 1),
                       "batched_p50": round(b["p50_ms"], 1), "batched_p99": round(b["p99_ms"], 1)},
        "dsm_median_abs_err_m": {"plain": p["dsm_err"], "batched": b["dsm_err"],
                                 "coverage_plain": round(p["dsm_cov"], 4),
                                 "coverage_batched": round(b["dsm_cov"], 4)},
        "roofline": roofline(b["stats"], cfg),
        "cpu_baseline": cpu,
        "fusion": {"gpu_ms_per_step": round(b["stats"]["fuse_ms"] / max(1, a.steps), 3)
                   if b["stats"]["fuse_ms"] else None,
                   "cpu_baseline": cpu_fuse},
    }
    if a.photometric_only:
        out["config"]["workload"] = (f"{cfg}: {desc}, {a.iters} iterations, r={a.window_radius}, "
                                     f"{res_m} m DSM fusion of the photometric maps (every cell kept)")
    ctx.close()
    return out


# ---------------------------------------------------------------------------
# serving
# ---------------------------------------------------------------------------
def bench_serving(a):
    from vissatsatellitestereo_amd import synth
    from vissatsatellitestereo_amd.serving import ServeLoop, poisson_arrivals
    works = []
    for k in range(4):
        spec = synth.spec_ragged(n_views=12, lo=1024, hi=a.serve_max_side, seed=50 + k, n_boxes=30,
                                 seed_offset=10 * (k + 1))
        rng = np.random.default_rng(60 + k)
        counts = [int(c) for c in rng.integers(8, 12, spec.n_views)]
        works.append(build_site(spec, os.path.join(a.work, f"c4_site{k}"), 11, counts))
    if a.build_only:
        return None
    from vissatsatellitestereo_amd import runtime
    sites = [load_site(w, id_base=1000 * (k + 1)) for k, w in enumerate(works)]
    grids = [runtime.dsm_grid(runtime.read_aoi(w)) for w in works]
    gts = [np.load(os.path.join(w, "synthetic_gt", "gt_dsm.npy")) for w in works]
    site_of = {v["ref"]: k for k, s in enumerate(sites) for v in s["views"]}
    copy_calibration(a)
    ctx = make_ctx(sites, a)
    views = [v for s in sites for v in s["views"]]
    ctx.set_params(num_iterations=a.iters, geom_consistency=0, filter=1)
    step, tune_rep = choose_step(ctx, views, a)
    rng = np.random.default_rng(6)
    order = rng.permutation(len(views))
    reqs = [views[i] for i in order][:a.requests]
    arrivals = poisson_arrivals(len(reqs), a.rate, seed=6)

    # incremental per-site DSM: each finished request re-fuses its site from
    # the photometric (filtered) depth maps of the site's finished requests
    fused = {"updates": 0, "ms": 0.0}
    finished = {k: [] for k in range(len(sites))}
    dsm = {}

    def on_complete(ticket, view, result):
        k = site_of[view["ref"]]
        finished[k].append(view["ref"])
        if len(finished[k]) >= 3:
            t = time.perf_counter()
            dsm[k] = ctx.fuse_dsm(grids[k], finished[k], depth_kind=0, min_measurements=3)
            fused["ms"] += 1e3 * (time.perf_counter() - t)
            fused["updates"] += 1

    def site_errors():
        errs = [dsm_error(dsm[k], gts[k])[0] for k in sorted(dsm)]
        errs = [e for e in errs if e is not None]
        return round(float(np.median(errs)), 4) if errs else None

    ctx.set_params(num_iterations=a.iters, geom_consistency=0, filter=1, window_step=step)
    warm = ServeLoop(ctx)                    # one request end to end
    warm.submit(reqs[0])
    warm.drain()
    loop = ServeLoop(ctx, on_complete=on_complete)
    ctx.reset_stats()
    served = loop.run_schedule(arrivals, reqs)
    st = ctx.stats()
    served_err = site_errors()
    log("served", {k: v for k, v in served.items() if k not in ("tickets", "latency_s")},
        "dsm", served_err, fused)
    # one-request-at-a-time queue on the same arrival schedule (plain form)
    ctx.set_params(window_step=1)
    t0 = time.perf_counter()
    lat = []
    for t_arr, v in zip(arrivals, reqs):
        now = time.perf_counter() - t0
        if now < t_arr:
            time.sleep(t_arr - now)
        ctx.solve([v], 0, want=("depth",))
        lat.append(time.perf_counter() - t0 - t_arr)
    plain_wall = time.perf_counter() - t0
    for k in list(dsm):
        dsm[k] = ctx.fuse_dsm(grids[k], finished[k], depth_kind=0, min_measurements=3)
    queue_err = site_errors()
    log(f"queue: {len(reqs) / plain_wall:.2f} views/s", "dsm", queue_err)
    # capacity: every request queued at t = 0, so neither form waits for
    # arrivals (at the configured rate the served form keeps up with the
    # arrivals and its views/s is the offered load)
    ctx.set_params(window_step=step)
    burst = ServeLoop(ctx).run_schedule([0.0] * len(reqs), reqs)
    ctx.set_params(window_step=1)
    ctx.synchronize()
    t0 = time.perf_counter()
    for v in reqs:
        ctx.solve([v], 0, want=("depth",))
    plain_cap = len(reqs) / (time.perf_counter() - t0)
    log(f"capacity: served {burst['views_per_s']:.2f} views/s, plain {plain_cap:.2f} views/s")
    out = {
        "metric": METRIC, "value": round(served["views_per_s"], 4), "unit": "views/s", "n_gpus": 1,
        "steps": len(reqs), "warmup": 1, "ms_per_step": round(1e3 * served["wall_s"] / len(reqs), 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"c4 serving: {len(reqs)} per-view requests, Poisson {a.rate} req/s, "
                               f"4 sites, crops 1024..{a.serve_max_side} per side, 8..11 sources, "
                               f"photometric {a.iters} iterations + filter, r={a.window_radius}, "
                               f"HBM budget {a.hbm_budget_gb} GB, incremental per-site 0.5 m DSM",
                   "forms": {"plain": "one request at a time in arrival order, window step 1",
                             "batched": f"continuous batching (vsm_submit / vsm_serve_step), "
                                        f"window step {step} (autotuner)"},
                   "autotune": tune_rep},
        "dsm_median_abs_err_m": {"plain": queue_err, "batched": served_err,
                                 "note": "median over sites of each site's median |dh|"},
        "incremental_fusion": {"updates": fused["updates"], "ms_total": round(fused["ms"], 1)},
        "latency_ms": {"served_p50": round(served["p50_ms"], 1), "served_p99": round(served["p99_ms"], 1),
                       "queue_p50": round(1e3 * float(np.percentile(lat, 50)), 1),
                       "queue_p99": round(1e3 * float(np.percentile(lat, 99)), 1)},
        "plain_views_per_s": round(len(reqs) / plain_wall, 4),
        "batched_over_plain": round(served["views_per_s"] / (len(reqs) / plain_wall), 3),
        "capacity": {"note": "all requests queued at t = 0",
                     "served_views_per_s": round(burst["views_per_s"], 4),
                     "plain_views_per_s": round(plain_cap, 4),
                     "served_over_plain": round(burst["views_per_s"] / plain_cap, 3),
                     "served_p50_ms": round(burst["p50_ms"], 1),
                     "served_p99_ms": round(burst["p99_ms"], 1)},
        "serve_steps": served["steps"],
        "roofline": roofline(st, "c4"),
        "cpu_baseline": None,
    }
    ctx.close()
    return out


def copy_calibration(a):
    """Optional HBM counter calibration: one elementwise kernel that reads and
    writes a_mb MiB each (dst = src * 1), after a flush buffer larger than the
    last-level cache has been written, so its reads come from HBM."""
    if a.copy_calibration_mb <= 0:
        return
    import torch
    n = a.copy_calibration_mb * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    flush = torch.ones(max(n, (1024 << 20) // 4), dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    log(f"copy calibration: {a.copy_calibration_mb} MiB read + written by one elementwise kernel")
    del src, dst, flush


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=("c1", "c2", "c3", "c4", "c5"))
    ap.add_argument("--iters", type=int, default=5)
    ap.add_argument("--window-radius", type=int, default=5)
    ap.add_argument("--batched-step", default="auto", help="auto (autotuner) or a fixed window step")
    ap.add_argument("--tune-steps", default="1,2",
                    help="window-step candidates of the autotuner (the batched configuration "
                         "allows strides 1 and 2; wider ones only on request)")
    ap.add_argument("--no-tune-cache", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--forms", default="plain,batched", help="comma list of plain,batched")
    ap.add_argument("--build-only", action="store_true", help="build the synthetic site(s) and exit")
    ap.add_argument("--copy-calibration-mb", type=int, default=0)
    ap.add_argument("--hbm-budget-gb", type=float, default=200.0)
    ap.add_argument("--rate", type=float, default=20.0, help="c4 arrival rate (req/s)")
    ap.add_argument("--requests", type=int, default=40, help="c4 number of requests")
    ap.add_argument("--serve-max-side", type=int, default=2048)
    ap.add_argument("--work", default=os.environ.get("VSM_BENCH_WORK", "/tmp/vsm_bench"))
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; runs on more GPUs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    out = bench_serving(a) if a.config == "c4" else bench_forms(a, a.config)
    if out is not None:
        print(json.dumps(out))


if __name__ == "__main__":
    main()
