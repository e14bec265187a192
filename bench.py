"""MVS throughput benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c5|c2|c3|c4|c1]
                    [--batched-step auto|1|2] [--tune-steps 1,2] [--no-cpu-baseline]
                    [--no-nested]

A step is one red-black half-step of every reference view of the workload
(images resident in HBM before timing).  A full solve is 2 x iterations
half-steps of the photometric pass followed by as many of the geometric-
consistency pass: with the default 5 + 5 iterations, 20 steps are exactly one
solve.  The photometric outputs are kept on the device in the step that ends
that pass, the geometric pass begins in the next step, and the filter, the
outputs and the 2.5D DSM fusion run in the solve's last step.  W warmup steps
run the same sequence on a solve that is discarded; the K timed steps start a
fresh solve (steps beyond 20 start the next one).  views/s = views x K / 20 /
wall time of the K steps.

Two forms on the same inputs:
  plain    one view per call (vsm_batch_* on a single view, synchronised),
           window step 1 (every tap, the reference's default schedule); a
           step runs that half-step for each view in turn;
  batched  all views in one persistent launch per half-step replayed from a
           captured hipGraph, window step chosen by the autotuner.
`value` = batched views/s.  Beside it: plain views/s, per-view latency
percentiles, the DSM median |error| of both forms against the synthetic
ground truth, the plain form at the batched form's window step
(`batched_over_plain_same_step`), the k_halfstep roofline and a CPU baseline.

Configurations (synthetic sites, seeded):
c5 (default)  40 views of 4096x4096 over a 2 km^2 site with 500 buildings,
              20 sources each, photometric + geometric 5 + 5 iterations,
              filter, 0.3 m DSM.  The 10 x 1024^2, 8-source workload (c2) is
              run after it and reported under "c2" (with the oracle accuracy
              bar of tests/golden/c2_oracle_dsm.json).
c2            10 views of 1024x1024, 8 sources, 0.5 m DSM.
c3            32 views of ragged crop sizes (512..2048 per side), 6..12 sources.
c4            serving: Poisson arrivals of per-view requests from 4 sites
              (17 views each, crops 1024..3072 per side, 8..16 sources), each
              request a photometric pass then the geometric pass with the
              filter; continuous batching with HBM-footprint admission,
              p50/p99 latency and views/s, against a one-solve-at-a-time
              queue on the same schedule; incremental per-site DSM.
c1            2 views of 256x256, 1 source, photometric only (the case the CPU
              oracle runs in full).
"""
import argparse
import json
import os
import platform
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
ORACLE_BAR_M = 0.05           # GPU DSM error may exceed the oracle's by at most this


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fp:
            for line in fp:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


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


def _sync(ctx):
    ctx.synchronize()
    import torch
    if torch.cuda.is_available():
        torch.cuda.synchronize()


# ---------------------------------------------------------------------------
# plain vs batched, step by step
# ---------------------------------------------------------------------------
PASS_PARAMS = {0: dict(geom_consistency=0, filter=0), 1: dict(geom_consistency=1, filter=1)}


class FormRunner:
    """The step sequence of one form over a set of views (see the module
    docstring): step k runs half-step k mod S of solve k // S for every view."""

    def __init__(self, ctx, views, form, window_step, iters, grid, photometric_only=False,
                 tile_width=32):
        self.ctx, self.views, self.form = ctx, list(views), form
        self.window_step, self.iters, self.grid = window_step, iters, grid
        self.tile_width = tile_width
        self.photometric_only = photometric_only
        self.P = 2 * iters
        self.S = self.P if photometric_only else 2 * self.P
        self.open = []                 # open batches (batched: one; plain: one per view)
        self.view_s = {}               # plain: per-view seconds of the current solve
        self.t_solve = None
        self.latencies = []            # per view (plain) / per solve, for every view (batched)
        self.dsm = None
        self.solves = 0

    def _params(self, pass_):
        self.ctx.set_params(num_iterations=self.iters, window_step=self.window_step,
                            tile_width=self.tile_width, **PASS_PARAMS[pass_])

    def _fuse(self):
        ids = [v["ref"] for v in self.views]
        if self.photometric_only:      # two views cannot pass the <= 2 rejection: keep every cell
            return self.ctx.fuse_dsm(self.grid, ids, depth_kind=0, min_measurements=1)
        return self.ctx.fuse_dsm(self.grid, ids, depth_kind=1, min_measurements=3)

    def step(self, k):
        j = k % self.S
        if self.form == "batched":
            self._step_batched(j)
        else:
            self._step_plain(j)

    def _step_batched(self, j):
        c = self.ctx
        if j == 0:
            self.t_solve = time.perf_counter()
            self._params(0)
            self.open = [c.batch(self.views, 0, use_graph=True)]
        elif j == self.P:
            self.open[0].end(want=())          # photometric maps stay on the device
            self._params(1)
            self.open = [c.batch(self.views, 1, use_graph=True)]
        self.open[0].step(1)
        if j == self.S - 1:
            self.open.pop().end(want=())
            lat = time.perf_counter() - self.t_solve
            self.latencies += [lat] * len(self.views)
            self.dsm = self._fuse()
            self.solves += 1

    def _timed(self, v, fn):
        t = time.perf_counter()
        r = fn()
        self.view_s[v["ref"]] += time.perf_counter() - t
        return r

    def _step_plain(self, j):
        c = self.ctx
        if j == 0:
            self.view_s = {v["ref"]: 0.0 for v in self.views}
            self._params(0)
            self.open = [self._timed(v, lambda v=v: c.batch([v], 0)) for v in self.views]
        elif j == self.P:
            for v, b in zip(self.views, self.open):
                self._timed(v, lambda b=b: b.end(want=()))
            self._params(1)
            self.open = [self._timed(v, lambda v=v: c.batch([v], 1)) for v in self.views]
        for v, b in zip(self.views, self.open):
            self._timed(v, lambda b=b: b.step(1))
        if j == self.S - 1:
            for v, b in zip(self.views, self.open):
                self._timed(v, lambda b=b: b.end(want=()))
            self.open = []
            self.latencies += list(self.view_s.values())
            self.dsm = self._fuse()
            self.solves += 1

    def discard(self):
        for b in self.open:
            b.discard()
        self.open = []


def run_form(ctx, views, form, window_step, a, grid, steps, warmup, photometric_only=False,
             timing=True, tile_width=32):
    """Warmup steps on a discarded solve, then `steps` timed steps from a
    fresh solve.  Returns per-step walls, the runner and the device stats of
    the timed steps."""
    r = FormRunner(ctx, views, form, window_step, a.iters, grid, photometric_only, tile_width)
    for k in range(warmup):
        r.step(k)
    r.discard()
    r = FormRunner(ctx, views, form, window_step, a.iters, grid, photometric_only, tile_width)
    ctx.reset_stats()
    ctx.set_timing(timing)
    walls = []
    _sync(ctx)
    for k in range(steps):
        t = time.perf_counter()
        r.step(k)
        walls.append(time.perf_counter() - t)
        if k % 5 == 4 or k == steps - 1:
            log(f"{form} (window step {window_step}) step {k}: {1e3 * walls[-1]:.0f} ms")
    _sync(ctx)
    st = ctx.stats()
    ctx.set_timing(False)
    r.discard()
    return walls, r, st


def summarize(walls, r, n_views, gt):
    tot = float(np.sum(walls))
    out = dict(ms_per_step=1e3 * float(np.mean(walls)),
               views_per_s=n_views * len(walls) / r.S / tot, wall_s=tot, solves=r.solves,
               step_ms=[round(1e3 * w, 2) for w in walls])
    if r.latencies:
        out["p50_ms"] = 1e3 * float(np.percentile(r.latencies, 50))
        out["p99_ms"] = 1e3 * float(np.percentile(r.latencies, 99))
    if r.dsm is not None:
        out["dsm_err"], out["dsm_cov"] = dsm_error(r.dsm, gt)
    return out


def roofline(st, cfg):
    if not (st["halfstep_ms"] > 0 and st["halfstep_launches"] > 0):
        return None
    avg_ms = st["halfstep_ms"] / st["halfstep_launches"]
    taps = st["taps"] / st["halfstep_launches"]
    ach = FLOPS_PER_TAP * taps / (avg_ms * 1e-3) / 1e12
    traffic, source, profiled = None, None, None
    # HBM traffic per launch comes from a committed counter profile of the
    # same configuration and kernel
    tfile = os.path.join(ROOT, "profiles", f"k_halfstep_{cfg}_counters.json")
    if os.path.exists(tfile):
        t = json.load(open(tfile))
        traffic = round(t["fetch_bytes"] + t["write_bytes"])
        source = (f"profiles/k_halfstep_{cfg}_counters.md: FETCH_SIZE + WRITE_SIZE per dispatch, "
                  "calibrated on an elementwise kernel of known bytes (batched form)"
                  + (f"; {t['calibration']}" if "calibration" in t else ""))
        gbps = traffic / (avg_ms * 1e-3) / 1e9
        profiled = {"hbm_gbps": round(gbps, 1), "hbm_frac": round(gbps / HBM_PEAK_GBPS, 4),
                    "valu_busy_pct": round(t["valubusy"], 1) if "valubusy" in t else None,
                    "valu_lanes_active_pct": round(t["valuutil"], 1) if "valuutil" in t else None,
                    "ta_busy_pct": round(t["ta_busy"], 1) if "ta_busy" in t else None,
                    "td_busy_pct": round(t["td_busy"], 1) if "td_busy" in t else None}
    return {"bound": "valu", "kernel": "k_halfstep", "achieved": round(ach, 3),
            "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(ach / FP32_PEAK_TFLOPS, 4),
            "traffic": traffic, "traffic_unit": "bytes per launch", "traffic_source": source,
            "profiled": profiled,
            "bound_note": "priced against vector FP32 (VALU), MI355X 157.3 TFLOP/s: a gather/stencil "
                          "kernel, neither MFMA nor HBM work; the profiled limiter is the texture "
                          "data path of its gathers (td_busy_pct) with VALU issue close behind",
            "avg_launch_ms": round(avg_ms, 4), "taps_per_launch": int(taps),
            "flops_per_tap": FLOPS_PER_TAP, "timing": "HIP events on the context stream around "
                                                      "every k_halfstep launch of the timed steps"}


def cpu_baseline(ctx, site, window_radius, iters, crop=512):
    """Oracle (FP64 C, OpenMP) on a bounded sample: a crop of the first
    reference view with all its sources, one photometric iteration, then one
    geometric iteration from that state against the sources' photometric
    depth maps the GPU left on the device; extrapolated to a full view (area)
    and to the workload's iterations of each pass."""
    import pm_oracle as po
    v = site["views"][0]
    inv_ids = {i: n for n, i in site["ids"].items()}
    ref = inv_ids[v["ref"]]
    srcs = [inv_ids[s] for s in v["srcs"]]
    imgs, P4 = site["imgs"], site["P4"]
    H, W = imgs[ref].shape
    ch, cw = min(crop, H), min(crop, W)
    y0, x0 = (H - ch) // 2, (W - cw) // 2
    crop_img = np.ascontiguousarray(imgs[ref][y0:y0 + ch, x0:x0 + cw])
    T = np.array([[1, 0, x0, 0], [0, 1, y0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], dtype=np.float64)
    P4c = np.linalg.inv(T) @ P4[ref]          # pixel (c, r) of the crop is (c + x0, r + y0)
    src_depths = [ctx.get_depth_map(s, kind=0) for s in v["srcs"]]
    ov = po.OracleView(crop_img, [imgs[s] for s in srcs], P4c, [P4[s] for s in srcs], v["depth_range"],
                       view_id=v["ref"], src_depths=src_depths,
                       params=dict(window_radius=window_radius))
    t = time.perf_counter()
    st = ov.solve(0, 1)
    t_photo = time.perf_counter() - t
    t = time.perf_counter()
    ov.solve(1, 1, state=st, do_filter=True)
    t_geo = time.perf_counter() - t
    scale = (H * W) / (ch * cw)
    per_view = scale * iters * (t_photo + t_geo)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(1.0 / per_view, 6), "unit": "views/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{ch}x{cw} crop of one {W}x{H} reference view, {len(srcs)} sources, window "
                      f"step 1: 1 photometric iteration of the FP64 oracle in {t_photo:.1f}s, then 1 "
                      f"geometric iteration + filter in {t_geo:.1f}s; per-view time = (sum) x "
                      f"{scale:.0f} (area) x {iters} (iterations per pass)"}


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
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"the whole workload: {n} views, {iters} photometric iterations, in {dt:.2f}s"}


def cpu_baseline_fusion(ctx, site, grid, max_views=2, band_rows=256):
    """numpy restatement of aggregate_2p5d (oracle/fusion_oracle.py), one
    process, on the geometric depth maps the batched form left on the device:
    per-view DSMs timed on up to max_views views and scaled to all of them,
    the cross-view fusion timed on a band of band_rows grid rows and scaled to
    the whole grid (the numpy median of big stacks runs for minutes)."""
    import fusion_oracle as fo
    ids = [v["ref"] for v in site["views"]]
    k = min(max_views, len(ids))
    t = time.perf_counter()
    per_view = []
    for i, n in zip(ids[:k], site["names"][:k]):
        per_view.append(fo.view_dsm(ctx.get_depth_map(i, kind=1), np.linalg.inv(site["P4"][n]), grid))
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
                     window_radius=a.window_radius, window_step=1, num_samples=15,
                     draw_group=a.draw_group)
    for s in sites:
        for n in s["names"]:
            ctx.add_image(s["ids"][n], s["imgs"][n], s["P4"][n])
    ctx.synchronize()
    return ctx


def choose_step(ctx, views, a):
    """Window step and tile width of the batched form: the autotuner's, or
    the command line's."""
    if a.batched_step != "auto":
        tw = 32 if a.tile_width == "auto" else int(a.tile_width)
        return int(a.batched_step), {"window_step": int(a.batched_step), "tile_width": tw,
                                     "source": "command line"}
    from vissatsatellitestereo_amd import autotune
    steps = tuple(int(x) for x in a.tune_steps.split(","))
    widths = autotune.TILE_WIDTHS if a.tile_width == "auto" else (int(a.tile_width),)
    st, rep = autotune.tune(ctx, views, iters=a.iters, radius=a.window_radius, steps=steps,
                            use_cache=not a.no_tune_cache, widths=widths)
    rep.setdefault("tile_width", widths[0])
    log("autotune", rep)
    return st, rep


def workload(cfg):
    from vissatsatellitestereo_amd import synth
    res_m, counts = 0.5, None
    if cfg == "c1":
        spec, max_src = synth.spec_c1(), 1
        desc = "2 views x 256x256 uint8, 1 source each, photometric only"
    elif cfg == "c2":
        spec, max_src = synth.spec_c2(), 8
        desc = "10 reference views x 1024x1024 uint8, 8 sources"
    elif cfg == "c5":
        spec, max_src = synth.spec_c5(), 20
        desc = "40 reference views x 4096x4096 uint8, 20 sources, 500 buildings"
        res_m = 0.3
    else:
        spec = synth.spec_ragged(32)
        rng = np.random.default_rng(7)
        counts = [int(k) for k in rng.integers(6, 13, spec.n_views)]
        max_src = 12
        desc = "32 reference views, ragged crops 512..2048 per side, 6..12 sources"
    return spec, max_src, counts, res_m, desc


def lds_staging_ab(a, cfg):
    """A/B of the window-tap gather path: every view's cost of its converged
    photometric depths (fronto-parallel planes) against every source, the taps
    read through the caches (A) or from a per-(tile, source) box of the source
    staged in LDS (B, where the box fits 4096 texels).  Same costs either way."""
    from vissatsatellitestereo_amd import runtime  # noqa: F401
    spec, max_src, counts, res_m, desc = workload(cfg)
    work = build_site(spec, os.path.join(a.work, cfg), max_src, counts, res_m)
    site = load_site(work)
    ctx = make_ctx([site], a)
    views = site["views"][:a.ab_views]
    out = {"workload": f"{cfg}: {desc}; {len(views)} views, r={a.window_radius}", "views": []}
    for ws in (1, 2):
        ctx.set_params(num_iterations=a.iters, window_step=ws, geom_consistency=0, filter=0)
        res = ctx.solve(views, 0, want=("depth",), use_graph=True)
        tot = {0: 0.0, 1: 0.0}
        for v, r in zip(views, res):
            n = np.zeros(r["depth"].shape + (3,), np.float32)
            n[..., 2] = 1.0
            c0, t0, _ = ctx.eval_cost_staging(v["ref"], v["srcs"], v["depth_range"], r["depth"], n, 0, a.ab_repeats)
            c1, t1, frac = ctx.eval_cost_staging(v["ref"], v["srcs"], v["depth_range"], r["depth"], n, 1,
                                                  a.ab_repeats)
            same = bool(np.array_equal(c0, c1))
            tot[0] += t0
            tot[1] += t1
            out["views"].append({"ref": v["ref"], "window_step": ws, "sources": len(v["srcs"]),
                                 "ms_global": round(t0, 3), "ms_lds": round(t1, 3),
                                 "staged_fraction": round(frac, 4), "identical_costs": same})
            log(out["views"][-1])
        out[f"step{ws}_lds_over_global_time"] = round(tot[1] / tot[0], 4)
    ctx.close()
    return out


def fusion_report(b, res, shapes, grid, cpu_fuse):
    """Device time of the 2.5D fusion per solve and its rate in algorithmic
    HBM bytes: 12 B per depth pixel (depth in, cell key and height out) and
    4 V + 4 B per DSM cell (the V-deep stack in, the fused cell out)."""
    if "batched" not in res or not b["stats"]["fuse_ms"]:
        return {"gpu_ms_per_solve": None, "cpu_baseline": cpu_fuse}
    ms = b["stats"]["fuse_ms"] / max(1, res["batched"]["solves"])
    cells = int(grid[4]) * int(grid[5])
    pixels = sum(int(h) * int(w) for h, w in shapes)
    nbytes = 12 * pixels + (4 * len(shapes) + 4) * cells
    gbps = nbytes / (ms * 1e-3) / 1e9
    return {"gpu_ms_per_solve": round(ms, 3), "algorithmic_bytes": nbytes,
            "achieved_gbps": round(gbps, 1), "hbm_frac": round(gbps / HBM_PEAK_GBPS, 4),
            "note": "a8-a12 chain on the device (per-view rasterise, hole fill, 3x3 median; "
                    "stack fusion; 3x3 median), timed with HIP events",
            "cpu_baseline": cpu_fuse}


def oracle_bar(cfg):
    f = os.path.join(ROOT, "tests", "golden", f"{cfg}_oracle_dsm.json")
    if not os.path.exists(f):
        return None
    return json.load(open(f))


def bench_forms(a, cfg, steps, warmup, with_cpu=True):
    from vissatsatellitestereo_amd import runtime
    photometric_only = cfg == "c1"
    spec, max_src, counts, res_m, desc = workload(cfg)
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
    res = {}
    tile_w = tune_rep["tile_width"]
    for form in [f for f in ("batched", "plain") if f in a.forms.split(",")]:
        ws = 1 if form == "plain" else step
        walls, r, st = run_form(ctx, views, form, ws, a, grid, steps, warmup, photometric_only,
                                tile_width=32 if form == "plain" else tile_w)
        res[form] = summarize(walls, r, len(views), gt)
        res[form]["stats"] = st
        log(form, {k: v for k, v in res[form].items() if k not in ("stats", "step_ms")})
    b = res.get("batched") or res["plain"]
    p = res.get("plain") or b
    # the plain form at the batched form's window step: what batching adds
    # without the stride.  On c5 over the first photometric half-steps only
    # (the timed region of the full leg takes ~100 s).
    same = None
    if "plain" in res and "batched" in res and step != 1 and not a.no_same_step:
        m = steps if cfg != "c5" else min(steps, a.same_step_steps)
        walls, r, _ = run_form(ctx, views, "plain", step, a, grid, m, min(warmup, 1), photometric_only,
                               timing=False, tile_width=tile_w)
        t_plain = float(np.sum(walls))
        t_batched = float(np.sum(b["step_ms"][:m])) / 1e3
        same = {"window_step": step, "steps": m,
                "plain_views_per_s": round(len(views) * m / r.S / t_plain, 4),
                "batched_over_plain": round(t_plain / t_batched, 3),
                "note": ("both forms on the first %d steps of a solve" % m) if m < steps
                        else "both forms on the whole timed region"}
        log("same window step", same)
    cpu = cpu_fuse = None
    if with_cpu and not a.no_cpu_baseline and "batched" in res:
        if cfg == "c1":
            cpu = cpu_baseline_full(site, a.window_radius, a.iters)
        else:
            # the sample reads the sources' photometric maps and the fusion
            # baseline the geometric maps of the last complete solve; with
            # fewer timed steps than one solve they come from one more batched
            # solve (outside every timed region)
            from vissatsatellitestereo_amd._lib import VsmError
            try:
                for v in views:
                    ctx.get_depth_map(v["ref"], kind=1)
            except VsmError:
                for pass_ in (0, 1):
                    ctx.set_params(num_iterations=a.iters, window_step=step, tile_width=tile_w,
                                   **PASS_PARAMS[pass_])
                    ctx.solve(views, pass_, want=(), use_graph=True)
            log("cpu baseline: PatchMatch oracle on a crop")
            cpu = cpu_baseline(ctx, site, a.window_radius, a.iters, crop=a.cpu_crop)
            log("cpu baseline: fusion")
            cpu_fuse = cpu_baseline_fusion(ctx, site, grid)
    ctx.close()
    errs = {"plain": p.get("dsm_err"), "batched": b.get("dsm_err"),
            "coverage_plain": round(p.get("dsm_cov", 0.0), 4),
            "coverage_batched": round(b.get("dsm_cov", 0.0), 4)}
    bar = oracle_bar(cfg)
    if bar is not None and errs["plain"] is not None and errs["batched"] is not None:
        lim = bar["dsm_median_abs_err_m"] + ORACLE_BAR_M
        errs["oracle"] = bar["dsm_median_abs_err_m"]
        errs["bar"] = round(lim, 5)
        errs["within_bar"] = bool(errs["plain"] <= lim and errs["batched"] <= lim)
        errs["oracle_source"] = "tests/golden/%s_oracle_dsm.json (%s)" % (cfg, bar.get("solver", ""))
    S = 2 * a.iters * (1 if photometric_only else 2)
    out = {
        "metric": METRIC, "value": round(b["views_per_s"], 4), "unit": "views/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": round(b["ms_per_step"], 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{cfg}: {desc}, photometric+geometric {a.iters}+{a.iters} iterations, "
                               f"filter, r={a.window_radius}, {res_m} m DSM fusion; a step is one "
                               f"half-step of every view, {S} steps one solve",
                   "forms": {"plain": "one view per call, window step 1, 32 x 16 tiles",
                             "batched": f"all views per launch, persistent kernel, hipGraph replay, "
                                        f"window step {step}, {tile_w} x {512 // tile_w} tiles "
                                        f"(autotuner)"},
                   "autotune": tune_rep,
                   "source_draws": ("per pixel group of one wave (tile width x 128 / tile width: "
                                    f"plain 32 x 4, batched {tile_w} x {128 // tile_w})"
                                    if a.draw_group else "per pixel")},
        "plain_views_per_s": round(p["views_per_s"], 4),
        "batched_over_plain": round(b["views_per_s"] / p["views_per_s"], 3),
        "batched_over_plain_same_step": same,
        "latency_ms": {"plain_p50": round(p.get("p50_ms", float("nan")), 1),
                       "plain_p99": round(p.get("p99_ms", float("nan")), 1),
                       "batched_p50": round(b.get("p50_ms", float("nan")), 1),
                       "batched_p99": round(b.get("p99_ms", float("nan")), 1)},
        "dsm_median_abs_err_m": errs,
        "roofline": roofline(b["stats"], cfg),
        "cpu_baseline": cpu,
        "fusion": fusion_report(b, res, [site["imgs"][n].shape for n in site["names"]], grid, cpu_fuse),
        "step_ms": {"batched": b["step_ms"], "plain": p["step_ms"]},
    }
    if photometric_only:
        out["config"]["workload"] = (f"{cfg}: {desc}, {a.iters} iterations, r={a.window_radius}, "
                                     f"{res_m} m DSM fusion of the photometric maps (every cell kept); "
                                     f"a step is one half-step of every view, {S} steps one solve")
    return out


# ---------------------------------------------------------------------------
# serving
# ---------------------------------------------------------------------------
def serving_sites(a):
    from vissatsatellitestereo_amd import synth
    works = []
    for k in range(4):
        spec = synth.spec_ragged(n_views=a.serve_views, lo=1024, hi=a.serve_max_side, seed=50 + k,
                                 n_boxes=30, seed_offset=10 * (k + 1))
        rng = np.random.default_rng(60 + k)
        counts = [int(c) for c in rng.integers(8, min(16, a.serve_views - 1) + 1, spec.n_views)]
        works.append(build_site(spec, os.path.join(a.work, f"c4_site{k}_{a.serve_views}_{a.serve_max_side}"),
                                16, counts))
    return works


def bench_serving(a):
    from vissatsatellitestereo_amd import runtime
    from vissatsatellitestereo_amd.serving import ServeLoop, poisson_arrivals
    works = serving_sites(a)
    if a.build_only:
        return None
    sites = [load_site(w, id_base=1000 * (k + 1)) for k, w in enumerate(works)]
    grids = [runtime.dsm_grid(runtime.read_aoi(w)) for w in works]
    gts = [np.load(os.path.join(w, "synthetic_gt", "gt_dsm.npy")) for w in works]
    site_of = {v["ref"]: k for k, s in enumerate(sites) for v in s["views"]}
    copy_calibration(a)
    ctx = make_ctx(sites, a)
    views = [v for s in sites for v in s["views"]]
    step, tune_rep = choose_step(ctx, views, a)
    rng = np.random.default_rng(6)
    order = rng.permutation(len(views))
    # every view of every site is requested once (a view's geometric pass
    # needs the photometric results of its sources, which are views of its site)
    reqs = [views[i] for i in order]
    arrivals = poisson_arrivals(len(reqs), a.rate, seed=6)
    n_src = [len(v["srcs"]) for v in reqs]
    sizes = [ctx.sizes[v["ref"]] for v in reqs]

    # incremental per-site DSM: each finished request re-fuses its site from
    # the geometric depth maps of the site's finished requests
    fused = {"updates": 0, "ms": 0.0}
    finished = {k: [] for k in range(len(sites))}
    dsm = {}

    def on_complete(ticket, view, result):
        k = site_of[view["ref"]]
        finished[k].append(view["ref"])
        if len(finished[k]) >= 3:
            t = time.perf_counter()
            dsm[k] = ctx.fuse_dsm(grids[k], finished[k], depth_kind=1, min_measurements=3)
            fused["ms"] += 1e3 * (time.perf_counter() - t)
            fused["updates"] += 1

    def site_errors():
        errs = [dsm_error(dsm[k], gts[k])[0] for k in sorted(dsm)]
        errs = [e for e in errs if e is not None]
        return round(float(np.median(errs)), 4) if errs else None

    tile_w = tune_rep["tile_width"]
    ctx.set_params(num_iterations=a.iters, window_step=step, tile_width=tile_w, **PASS_PARAMS[1])
    warm = ServeLoop(ctx)                              # two photometric requests end to end
    for v in sites[0]["views"][:2]:
        warm.submit(v)
    warm.drain()
    # results stay on the device (the per-site fusion reads them there), as
    # in the one-at-a-time queue below: neither form copies maps to the host
    loop = ServeLoop(ctx, on_complete=on_complete, geometric=True, want=())
    ctx.reset_stats()
    ctx.set_timing(True)
    served = loop.run_schedule(arrivals, reqs)
    st = ctx.stats()
    ctx.set_timing(False)
    served_err = site_errors()
    log("served", {k: v for k, v in served.items() if k not in ("tickets", "latency_s")},
        "dsm", served_err, fused)
    # plain: one solve at a time in arrival order, each a single view at
    # window step 1; a view's geometric solve runs once the photometric
    # solves of the view and its sources are done
    ctx.set_params(window_step=1, tile_width=32)
    plain = queue_schedule(ctx, arrivals, reqs)
    for k in list(dsm):
        dsm[k] = ctx.fuse_dsm(grids[k], finished[k], depth_kind=1, min_measurements=3)
    queue_err = site_errors()
    log(f"queue: {plain['views_per_s']:.2f} views/s", "dsm", queue_err)
    # capacity: every request queued at t = 0 (at the configured rate the
    # served form keeps up with the arrivals and its views/s is the offered load)
    ctx.set_params(window_step=step, tile_width=tile_w)
    burst = ServeLoop(ctx, geometric=True, want=()).run_schedule([0.0] * len(reqs), reqs)
    ctx.set_params(window_step=1, tile_width=32)
    plain_cap = queue_schedule(ctx, [0.0] * len(reqs), reqs)
    log(f"capacity: served {burst['views_per_s']:.2f} views/s, plain {plain_cap['views_per_s']:.2f} views/s")
    out = {
        "metric": METRIC, "value": round(served["views_per_s"], 4), "unit": "views/s", "n_gpus": 1,
        "steps": len(reqs), "warmup": 1, "ms_per_step": round(1e3 * served["wall_s"] / len(reqs), 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"c4 serving: {len(reqs)} per-view requests, Poisson {a.rate} req/s, "
                               f"4 sites x {a.serve_views} views, crops 1024..{a.serve_max_side} per side "
                               f"(largest {max(max(s) for s in sizes)}), {min(n_src)}..{max(n_src)} "
                               f"sources; each request photometric {a.iters} iterations then geometric "
                               f"{a.iters} iterations + filter, r={a.window_radius}, HBM budget "
                               f"{a.hbm_budget_gb} GB, incremental per-site 0.5 m DSM of the geometric maps",
                   "forms": {"plain": "one solve at a time in arrival order, window step 1",
                             "batched": f"continuous batching (vsm_submit_pass / vsm_serve_step), "
                                        f"window step {step} (autotuner)"},
                   "autotune": tune_rep},
        "dsm_median_abs_err_m": {"plain": queue_err, "batched": served_err,
                                 "note": "median over sites of each site's median |dh|"},
        "incremental_fusion": {"updates": fused["updates"], "ms_total": round(fused["ms"], 1)},
        "latency_ms": {"served_p50": round(served["p50_ms"], 1), "served_p99": round(served["p99_ms"], 1),
                       "queue_p50": round(plain["p50_ms"], 1), "queue_p99": round(plain["p99_ms"], 1)},
        "plain_views_per_s": round(plain["views_per_s"], 4),
        "batched_over_plain": round(served["views_per_s"] / plain["views_per_s"], 3),
        "capacity": {"note": "all requests queued at t = 0",
                     "served_views_per_s": round(burst["views_per_s"], 4),
                     "plain_views_per_s": round(plain_cap["views_per_s"], 4),
                     "served_over_plain": round(burst["views_per_s"] / plain_cap["views_per_s"], 3),
                     "served_p50_ms": round(burst["p50_ms"], 1),
                     "served_p99_ms": round(burst["p99_ms"], 1)},
        "serve_steps": served["steps"],
        "served_gpu_halfstep_s": round(st["halfstep_ms"] / 1e3, 3),
        "hbm_peak_gb": round(st["hbm_bytes_peak"] / 1e9, 2),
        "roofline": roofline(st, "c4"),
        "cpu_baseline": None,
    }
    ctx.close()
    return out


def queue_schedule(ctx, arrivals, reqs):
    """One solve at a time in arrival order: a view's photometric solve when
    it arrives, its geometric solve (with the filter) as soon as the
    photometric solves of the view and all its sources are done.  Latency of
    a request = its geometric solve's end - its arrival."""
    order = np.argsort(arrivals)
    photo_done, geo_done = set(), {}
    t0 = time.perf_counter()
    waiting = []
    for i in order:
        v = reqs[int(i)]
        now = time.perf_counter() - t0
        if now < arrivals[i]:
            time.sleep(arrivals[i] - now)
        ctx.set_params(**PASS_PARAMS[0])
        ctx.solve([v], 0, want=())
        photo_done.add(v["ref"])
        waiting.append(int(i))
        for j in list(waiting):
            w = reqs[j]
            if w["ref"] in photo_done and all(s in photo_done for s in w["srcs"]):
                ctx.set_params(**PASS_PARAMS[1])
                ctx.solve([w], 1, want=())
                geo_done[j] = time.perf_counter() - t0
                waiting.remove(j)
    wall = time.perf_counter() - t0
    lat = [geo_done[j] - arrivals[j] for j in geo_done]
    return dict(views_per_s=len(geo_done) / wall, wall_s=wall, completed=len(geo_done),
                p50_ms=1e3 * float(np.percentile(lat, 50)), p99_ms=1e3 * float(np.percentile(lat, 99)))


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
    ap.add_argument("--steps", type=int, default=20, help="timed steps (a step = one half-step of "
                                                          "every view; 20 = one 5+5 solve)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c5", choices=("c1", "c2", "c3", "c4", "c5"))
    ap.add_argument("--iters", type=int, default=5)
    ap.add_argument("--window-radius", type=int, default=5)
    ap.add_argument("--batched-step", default="auto", help="auto (autotuner) or a fixed window step")
    ap.add_argument("--tune-steps", default="1,2",
                    help="window-step candidates of the autotuner (the batched configuration "
                         "allows strides 1 and 2; wider ones only on request)")
    ap.add_argument("--no-tune-cache", action="store_true")
    ap.add_argument("--tile-width", default="auto", help="auto (autotuner) or 16, 32, 64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-same-step", action="store_true")
    ap.add_argument("--same-step-steps", type=int, default=4,
                    help="c5: steps of the plain form at the batched window step")
    ap.add_argument("--no-nested", action="store_true", help="c5: skip the nested c2 workload")
    ap.add_argument("--cpu-crop", type=int, default=512)
    ap.add_argument("--draw-group", type=int, default=1, choices=(0, 1),
                    help="Monte-Carlo source draws keyed by pixel (0) or by the pixel group one "
                         "wave covers (1, tile width x 128 / tile width; both forms)")
    ap.add_argument("--forms", default="plain,batched", help="comma list of plain,batched")
    ap.add_argument("--build-only", action="store_true", help="build the synthetic site(s) and exit")
    ap.add_argument("--copy-calibration-mb", type=int, default=0)
    ap.add_argument("--hbm-budget-gb", type=float, default=200.0)
    ap.add_argument("--rate", type=float, default=20.0, help="c4 arrival rate (req/s)")
    ap.add_argument("--serve-views", type=int, default=17, help="c4 views per site")
    ap.add_argument("--serve-max-side", type=int, default=3072)
    ap.add_argument("--work", default=os.environ.get("VSM_BENCH_WORK", "/tmp/vsm_bench"))
    ap.add_argument("--lds-staging-ab", action="store_true",
                    help="measure LDS-staged against cached gathers of source footprints and exit")
    ap.add_argument("--ab-views", type=int, default=4)
    ap.add_argument("--ab-repeats", type=int, default=3)
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; runs on more GPUs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.lds_staging_ab:
        print(json.dumps(lds_staging_ab(a, a.config)))
        return
    if a.config == "c4":
        out = bench_serving(a)
    else:
        out = bench_forms(a, a.config, a.steps, a.warmup)
        if out is not None and a.config == "c5" and not a.no_nested:
            log("nested workload: c2")
            nested = bench_forms(a, "c2", 20, min(a.warmup, 5), with_cpu=False)
            keep = ("value", "ms_per_step", "plain_views_per_s", "batched_over_plain",
                    "batched_over_plain_same_step", "latency_ms", "dsm_median_abs_err_m", "roofline")
            out["c2"] = {k: nested[k] for k in keep}
            out["c2"]["workload"] = nested["config"]["workload"]
            out["c2"]["window_step"] = nested["config"]["autotune"].get("window_step")
    if out is not None:
        print(json.dumps(out))


if __name__ == "__main__":
    main()
