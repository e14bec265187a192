"""bench.py's step sequence on the CPU with a fake context: a step is one
half-step of every view, 2 x iterations steps per pass, the photometric
outputs are kept before the geometric pass begins, and the filter/outputs
and the fusion run in the solve's last step."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


class FakeBatch:
    def __init__(self, log, views, pass_, graph, nsteps):
        self.log, self.views, self.pass_, self.graph = log, views, pass_, graph
        self.left = nsteps
        log.append(("begin", tuple(v["ref"] for v in views), pass_, graph))

    def step(self, n=1):
        assert self.left > 0
        self.left -= n
        self.log.append(("step", tuple(v["ref"] for v in self.views), self.pass_))
        return self.left

    def end(self, want=()):
        assert self.left == 0
        self.log.append(("end", tuple(v["ref"] for v in self.views), self.pass_))
        return [{} for _ in self.views]

    def discard(self):
        self.log.append(("discard", tuple(v["ref"] for v in self.views), self.pass_))


class FakeCtx:
    def __init__(self):
        self.log, self.params = [], {}

    def set_params(self, **kw):
        self.params.update(kw)

    def batch(self, views, pass_=0, use_graph=False):
        assert self.params["geom_consistency"] == pass_ and self.params["filter"] == pass_
        return FakeBatch(self.log, views, pass_, use_graph, 2 * self.params["num_iterations"])

    def fuse_dsm(self, grid, ids, depth_kind=1, min_measurements=3):
        self.log.append(("fuse", tuple(ids), depth_kind, min_measurements))
        return None


VIEWS = [dict(ref=r, srcs=[], depth_range=(0, 1)) for r in (1, 2, 3)]


def test_batched_steps_make_one_solve_in_20():
    ctx = FakeCtx()
    r = bench.FormRunner(ctx, VIEWS, "batched", 2, 5, None)
    for k in range(20):
        r.step(k)
    kinds = [e[0] for e in ctx.log]
    assert kinds.count("step") == 20 and kinds.count("fuse") == 1
    begins = [e for e in ctx.log if e[0] == "begin"]
    assert [b[2] for b in begins] == [0, 1] and all(b[3] for b in begins) and all(len(b[1]) == 3 for b in begins)
    # the photometric pass ends after its 10th half-step, before the geometric begin
    i_end0 = ctx.log.index(("end", (1, 2, 3), 0))
    assert sum(1 for e in ctx.log[:i_end0] if e[0] == "step") == 10
    assert ctx.log.index(begins[1]) == i_end0 + 1
    assert ctx.log[-2] == ("end", (1, 2, 3), 1) and ctx.log[-1] == ("fuse", (1, 2, 3), 1, 3)
    assert r.solves == 1 and len(r.latencies) == 3 and ctx.params["window_step"] == 2


def test_plain_steps_run_each_view_alone():
    ctx = FakeCtx()
    r = bench.FormRunner(ctx, VIEWS, "plain", 1, 5, None)
    for k in range(20):
        r.step(k)
    steps = [e for e in ctx.log if e[0] == "step"]
    assert len(steps) == 60 and all(len(e[1]) == 1 for e in steps)
    # every photometric batch ends before the first geometric one begins
    last_end0 = max(i for i, e in enumerate(ctx.log) if e[0] == "end" and e[2] == 0)
    first_begin1 = min(i for i, e in enumerate(ctx.log) if e[0] == "begin" and e[2] == 1)
    assert last_end0 < first_begin1
    assert ctx.log[-1] == ("fuse", (1, 2, 3), 1, 3)
    assert r.solves == 1 and len(r.latencies) == 3 and ctx.params["window_step"] == 1


def test_warmup_discards_open_batches_and_photometric_only():
    ctx = FakeCtx()
    r = bench.FormRunner(ctx, VIEWS, "batched", 2, 5, None)
    for k in range(5):
        r.step(k)
    r.discard()
    assert ctx.log[-1][0] == "discard" and r.open == []
    ctx = FakeCtx()
    r = bench.FormRunner(ctx, VIEWS[:2], "batched", 2, 5, None, photometric_only=True)
    for k in range(10):
        r.step(k)
    assert r.S == 10 and ctx.log[-1] == ("fuse", (1, 2), 0, 1)
    assert not any(e[0] == "begin" and e[2] == 1 for e in ctx.log)


def test_c2_oracle_fixture_and_bar():
    """The committed oracle fixture for the 10 x 1024^2 workload exists, was
    made by the FP64 oracle over the full pipeline at stride 1, and the
    benchmark's accuracy bar is its error + 0.05 m."""
    bar = bench.oracle_bar("c2")
    assert bar is not None
    assert "stride 1" in bar["workload"] and "5+5" in bar["workload"]
    assert "pm_oracle.c" in bar["solver"]
    assert 0.0 < bar["dsm_median_abs_err_m"] < 0.1
    assert bar["dsm_coverage"] > 0.9 and bar["kept_fraction"] > 0.5
    assert bench.ORACLE_BAR_M == 0.05
    assert bench.oracle_bar("c9") is None


def test_roofline_from_live_timing_and_committed_profile():
    """roofline(): achieved = 30 FLOP x device-counted taps / live launch time;
    traffic and the busy figures come from the committed counter profile of
    the same workload (none for a workload without one)."""
    import json
    st = {"halfstep_ms": 2 * 1250.0, "halfstep_launches": 2, "taps": 2 * 1.104e12}
    r = bench.roofline(st, "c5")
    assert r["kernel"] == "k_halfstep" and r["unit"] == "TFLOP/s"
    assert abs(r["achieved"] - 30 * 1.104e12 / 1.25 / 1e12) < 1e-3
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-4 and r["peak"] == 157.3
    prof = json.load(open(os.path.join(ROOT, "profiles", "k_halfstep_c5_counters.json")))
    assert r["traffic"] == round(prof["fetch_bytes"] + prof["write_bytes"])
    p = r["profiled"]
    assert p["td_busy_pct"] == round(prof["td_busy"], 1) and p["ta_busy_pct"] == round(prof["ta_busy"], 1)
    assert abs(p["hbm_gbps"] - r["traffic"] / 1.25 / 1e9) < 0.1
    assert bench.roofline(st, "c9")["traffic"] is None
    assert bench.roofline({"halfstep_ms": 0.0, "halfstep_launches": 0, "taps": 0}, "c5") is None
