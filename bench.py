"""MVS throughput benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2]
                    [--batched-step 1|2] [--no-cpu-baseline]

Workload (default, config c2): a synthetic site of 10 views of 1024x1024
uint8 (5-30 deg off-nadir, 8 azimuths, 600 km cameras); every view is a
reference with 8 sources; photometric pass (5 iterations) then geometric
pass with filter (5 iterations); then 2.5D fusion into a 0.5 m DSM.  One
"step" solves all 10 reference views and fuses them, with the images already
resident in HBM.

Two forms are timed on the same inputs:
  plain    one reference view per vsm_solve call, synchronised per view;
  batched  all views in one vsm_solve per pass, the half-step sequence
           replayed from a captured hipGraph.
`value` is the batched views/s; the plain figure, per-view latency
percentiles and the DSM error against the synthetic ground truth of both
forms are reported beside it.  One JSON line on stdout.
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
FLOPS_PER_TAP = 30            # homography 12, divide 3, bilinear 8, accumulate 7


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_site(cfg, work):
    from vissatsatellitestereo_amd import synth, workspace
    import torch
    dev = "cuda" if torch.cuda.is_available() else "cpu"
    if cfg == "c2":
        spec = synth.spec_c2()
        n_src = 8
    elif cfg == "c2s":       # reduced c2 for quick checks
        spec = synth.spec_c2(n_views=6, size=512)
        n_src = 5
    else:
        raise SystemExit(f"unknown config {cfg}")
    if not os.path.exists(os.path.join(work, "colmap", "mvs", "depth_ranges.txt")):
        t = time.time()
        synth.build_site(spec, work, device=dev, save_gt_views=False, log=log)
        workspace.prepare_dense_workspace(work, max_src=n_src)
        log(f"site {cfg} built in {time.time() - t:.1f}s")
    return spec, n_src


def load_views(work):
    from vissatsatellitestereo_amd import reparam_depth as rd
    from vissatsatellitestereo_amd.workspace import select_sources
    from PIL import Image
    mvs = os.path.join(work, "colmap", "mvs")
    P4 = rd.projection_matrices(mvs)
    names = sorted(P4)
    imgs = {n: np.array(Image.open(os.path.join(mvs, "images", n))) for n in names}
    dr = rd.read_name_values(os.path.join(mvs, "depth_ranges.txt"))
    srcs = select_sources(mvs)
    ids = {n: i + 1 for i, n in enumerate(names)}
    views = [dict(ref=ids[n], srcs=[ids[s] for s in srcs[n]], depth_range=dr[n]) for n in names]
    return names, imgs, P4, views, ids


def dsm_error(dsm, gt):
    ok = ~np.isnan(dsm) & ~np.isnan(gt)
    if not ok.any():
        return None, 0.0
    return float(np.median(np.abs(dsm[ok] - gt[ok]))), float(ok.mean())


def run_form(ctx, views, grid, form, iters, use_graph):
    """One step: photometric + geometric for every view, then fusion.
    Returns (wall seconds, per-view latencies, fused DSM)."""
    lat = {}
    t0 = time.perf_counter()
    for pass_, extra in ((0, dict(geom_consistency=0, filter=0)),
                         (1, dict(geom_consistency=1, filter=1))):
        ctx.set_params(num_iterations=iters, **extra)
        if form == "plain":
            for v in views:
                ts = time.perf_counter()
                ctx.solve([v], pass_, want=("depth",))
                lat[v["ref"]] = lat.get(v["ref"], 0.0) + time.perf_counter() - ts
        else:
            ctx.solve(views, pass_, want=(), use_graph=use_graph)
    dsm = ctx.fuse_dsm(grid, [v["ref"] for v in views], depth_kind=1, min_measurements=3)
    wall = time.perf_counter() - t0
    if form != "plain":
        lat = {v["ref"]: wall for v in views}
    return wall, list(lat.values()), dsm


def cpu_baseline(names, imgs, P4, views, ids, window_radius, seconds_hint=20.0):
    """Oracle (FP64 C, OpenMP) on a bounded sample: a 512x512 crop of the
    first reference view with all its sources, 1 photometric iteration;
    extrapolated to a full view (area x iterations)."""
    import pm_oracle as po
    v = views[0]
    inv_ids = {i: n for n, i in ids.items()}
    ref = inv_ids[v["ref"]]
    srcs = [inv_ids[s] for s in v["srcs"]]
    H, W = imgs[ref].shape
    ch, cw = min(512, H), min(512, W)
    y0, x0 = (H - ch) // 2, (W - cw) // 2
    crop = np.ascontiguousarray(imgs[ref][y0:y0 + ch, x0:x0 + cw])
    T = np.array([[1, 0, x0, 0], [0, 1, y0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], dtype=np.float64)
    # P4 of the crop: pixel (c, r) of the crop is (c + x0, r + y0) of the view
    P4c = np.linalg.inv(T) @ P4[ref]
    ov = po.OracleView(crop, [imgs[s] for s in srcs], P4c, [P4[s] for s in srcs], v["depth_range"],
                       view_id=v["ref"], params=dict(window_radius=window_radius))
    t = time.perf_counter()
    ov.solve(0, 1)
    dt = time.perf_counter() - t
    scale = (H * W) / (ch * cw)
    return dt, scale, f"{ch}x{cw} crop of one {W}x{H} reference view, {len(srcs)} sources, " \
                      f"1 photometric iteration ({dt:.1f}s); per-view time = sample x {scale:.0f} " \
                      f"(area) x iterations"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--iters", type=int, default=5)
    ap.add_argument("--window-radius", type=int, default=5)
    ap.add_argument("--batched-step", type=int, default=1, help="window stride of the batched form")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--forms", default="plain,batched", help="comma list of plain,batched")
    ap.add_argument("--build-only", action="store_true", help="build the synthetic site and exit")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="run a device copy of this many MiB first (HBM counter calibration)")
    ap.add_argument("--work", default=os.environ.get("VSM_BENCH_WORK", "/tmp/vsm_bench"))
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)

    from vissatsatellitestereo_amd.mvs import MvsContext
    work = os.path.join(a.work, a.config)
    spec, n_src = build_site(a.config, work)
    if a.build_only:
        return
    if a.copy_calibration_mb > 0:
        import torch
        n = a.copy_calibration_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {a.copy_calibration_mb} MiB read + written")
        del src, dst
    names, imgs, P4, views, ids = load_views(work)
    gt = np.load(os.path.join(work, "synthetic_gt", "gt_dsm.npy"))
    from vissatsatellitestereo_amd import runtime
    grid = runtime.dsm_grid(runtime.read_aoi(work))

    ctx = MvsContext(device=0, window_radius=a.window_radius, window_step=1, num_samples=15)
    for n in names:
        ctx.add_image(ids[n], imgs[n], P4[n])
    ctx.synchronize()
    H, W = imgs[names[0]].shape
    nviews = len(views)

    res = {}
    forms = [f for f in ("plain", "batched") if f in a.forms.split(",")]
    for form in forms:
        ctx.set_params(window_step=1 if form == "plain" else a.batched_step)
        use_graph = form == "batched"
        for _ in range(a.warmup):
            run_form(ctx, views, grid, form, a.iters, use_graph)
        log(f"{form}: warm")
        ctx.reset_stats()
        ctx.set_timing(form == "batched")
        walls, lats = [], []
        dsm = None
        ctx.synchronize()
        for k in range(a.steps):
            w, lat, dsm = run_form(ctx, views, grid, form, a.iters, use_graph)
            walls.append(w)
            lats += lat
            log(f"{form} step {k}: {w * 1e3:.0f} ms")
        ctx.synchronize()
        st = ctx.stats()
        ctx.set_timing(False)
        err, cov = dsm_error(dsm, gt)
        res[form] = dict(wall=float(np.sum(walls)), ms_per_step=1e3 * float(np.mean(walls)),
                         views_per_s=nviews * a.steps / float(np.sum(walls)),
                         p50_ms=1e3 * float(np.percentile(lats, 50)),
                         p99_ms=1e3 * float(np.percentile(lats, 99)),
                         dsm_err=err, dsm_cov=cov, stats=st)
        log(form, {k: v for k, v in res[form].items() if k != "stats"})

    b = res.get("batched") or res["plain"]
    st = b["stats"]
    roof = None
    if st["halfstep_ms"] > 0 and st["halfstep_launches"] > 0:
        avg_ms = st["halfstep_ms"] / st["halfstep_launches"]
        taps_per_launch = st["taps"] / st["halfstep_launches"]
        ach = FLOPS_PER_TAP * taps_per_launch / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "valu", "kernel": "k_halfstep", "achieved": round(ach, 3),
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(ach / FP32_PEAK_TFLOPS, 4),
                "traffic": None, "avg_launch_ms": round(avg_ms, 4),
                "taps_per_launch": int(taps_per_launch), "flops_per_tap": FLOPS_PER_TAP}

    cpu = None
    if not a.no_cpu_baseline:
        try:
            dt, scale, sample = cpu_baseline(names, imgs, P4, views, ids, a.window_radius)
            per_view = dt * scale * (2 * a.iters)
            cpu = {"value": round(1.0 / per_view, 6), "unit": "views/s",
                   "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                   "kind": "port", "sample": sample}
        except Exception as e:      # the baseline is reported, never required
            cpu = {"value": None, "unit": "views/s", "cores": None, "kind": "port",
                   "sample": f"failed: {e}"}

    p = res.get("plain") or b
    out = {
        "metric": METRIC, "value": round(b["views_per_s"], 4), "unit": "views/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(b["ms_per_step"], 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{a.config}: {nviews} reference views x {W}x{H} uint8, {n_src} sources, "
                               f"photometric+geometric {a.iters}+{a.iters} iterations, filter, "
                               f"r={a.window_radius}, 0.5 m DSM fusion",
                   "forms": {"plain": "one view per solve, window step 1",
                             "batched": f"all views per solve, hipGraph replay, window step {a.batched_step}"}},
        "plain_views_per_s": round(p["views_per_s"], 4),
        "batched_over_plain": round(b["views_per_s"] / p["views_per_s"], 3),
        "latency_ms": {"plain_p50": round(p["p50_ms"], 1), "plain_p99": round(p["p99_ms"], 1),
                       "batched_p50": round(b["p50_ms"], 1), "batched_p99": round(b["p99_ms"], 1)},
        "dsm_median_abs_err_m": {"plain": p["dsm_err"], "batched": b["dsm_err"],
                                 "coverage_plain": round(p["dsm_cov"], 4),
                                 "coverage_batched": round(b["dsm_cov"], 4)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
