"""The HIP path at the benchmark workloads' real shapes.

* c2: 1024^2 views with 8 sources -- cost of fixed hypotheses against the
  oracle at full-image coordinates, lockstep half-steps (strides 1 and 2) on
  a crop of a real view against its full-size sources, batched == single for
  both passes at the full size, and the DSM error of a full solve against the
  synthetic ground truth and the oracle's (tests/golden/c2_oracle_dsm.json);
* c3: 32 ragged views 512..2048 with 6..12 sources -- one graph-replayed
  batch bit-equal to one-view solves;
* c4: crops up to 3072^2 with up to 16 sources -- served alone == mixed
  under a budget that forces queueing, and the served geometric pass ==
  batched;
* c5: 4096^2 views with 20 sources (the kernel for more than 16 sources) --
  cost against the oracle at window step 1 on a crop at full-image
  coordinates, two lockstep half-steps at step 1, batched == single at the
  full grid, and 0.3 m fusion bit-exact against the numpy oracle.

Oracle checks run on crops: a crop of the reference with the projection
P4c = inv(T) P4 (T shifts pixel (c, r) to (c + x0, r + y0)) is the same
reference geometry, so the oracle and the GPU evaluate identical windows away
from the crop border.
"""
import json
import os

import numpy as np
import pytest

import sites

pytestmark = pytest.mark.gpu

COST_TOL = 1e-4


def _params(**kw):
    """Solver parameters shared by the GPU context and the oracle (the oracle
    takes its depth range from the view)."""
    p = dict(window_radius=5, window_step=1, num_iterations=5, num_samples=15)
    p.update(kw)
    return p


GPU_KEYS = ("window_radius", "window_step", "num_samples", "num_iterations", "sigma_color",
            "ncc_sigma", "min_triangulation_angle", "incident_angle_sigma",
            "geom_consistency_regularizer", "geom_consistency_max_cost", "filter_min_ncc",
            "filter_min_triangulation_angle", "filter_geom_consistency_max_cost",
            "filter_min_num_consistent", "perturbation", "normal_perturbation", "max_normal_angle",
            "perturb_decay", "n_far", "seed", "draw_group", "tile_width")


def _ctx(site, p, ids=None, budget=200 * 10**9):
    import pm_oracle as po
    from vissatsatellitestereo_amd.mvs import MvsContext
    full = po.default_params(**p)
    ctx = MvsContext(device=0, hbm_budget_bytes=budget, **{k: full[k] for k in GPU_KEYS if k in full})
    for n in (ids or site["names"]):
        ctx.add_image(_id(site, n), site["imgs"][n], site["P4"][n])
    return ctx


def _id(site, name):
    return site["names"].index(name) + 1


def _sources(site, ref, k):
    from vissatsatellitestereo_amd.workspace import select_sources
    return select_sources(site["mvs"])[ref][:k]


def _crop(site, ref, y0, x0, h, w):
    img = np.ascontiguousarray(site["imgs"][ref][y0:y0 + h, x0:x0 + w])
    T = np.array([[1, 0, x0, 0], [0, 1, y0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], dtype=np.float64)
    return img, np.linalg.inv(T) @ site["P4"][ref]


def _oracle(site, ref_img, P4_ref, srcs, dr, view_id, p, src_depths=None):
    import pm_oracle as po
    return po.OracleView(ref_img, [site["imgs"][s] for s in srcs], P4_ref, [site["P4"][s] for s in srcs],
                         dr, view_id=view_id, src_depths=src_depths, params=p)


def _hyps(shape, dr, seed, good_depth=None):
    """Random planes; where good_depth is given, half the rows carry it with a
    fronto-parallel normal (well-matched windows, NCC near 1)."""
    rng = np.random.default_rng(seed)
    d = rng.uniform(dr[0], dr[1], shape).astype(np.float32)
    n = rng.normal(size=shape + (3,)) * [0.3, 0.3, 1.0]
    n[..., 2] = np.abs(n[..., 2]) + 0.3
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    n = n.astype(np.float32)
    if good_depth is not None:
        d[::2] = good_depth[::2]
        n[::2] = [0, 0, 1]
    return d, n


def _cost_vs_oracle(site, ref, srcs, p, y0, x0, h, w, good=None, seed=3):
    """GPU cost of fixed hypotheses over the full reference view against the
    oracle's on a crop, compared away from the crop border."""
    ctx = _ctx(site, p, [ref] + srcs)
    try:
        H, W = site["imgs"][ref].shape
        d, n = _hyps((H, W), site["dr"][ref], seed, good)
        g = ctx.eval_cost(_id(site, ref), [_id(site, s) for s in srcs], site["dr"][ref], d, n, 0)
    finally:
        ctx.close()
    img, P4c = _crop(site, ref, y0, x0, h, w)
    ov = _oracle(site, img, P4c, srcs, site["dr"][ref], _id(site, ref), p)
    o = ov.cost_map(d[y0:y0 + h, x0:x0 + w].astype(np.float64), n[y0:y0 + h, x0:x0 + w].astype(np.float64))
    r = p["window_radius"]
    gg = g[:, y0 + r:y0 + h - r, x0 + r:x0 + w - r]
    oo = o[:, r:h - r, r:w - r]
    err = np.abs(gg - oo)
    print(f"cost vs oracle ({len(srcs)} sources, step {p['window_step']}): max {err.max():.2e}, "
          f"matched windows (cost < 0.3) {np.mean(oo < 0.3):.3f}")
    return err, oo


def _lockstep_crop(site, ref, srcs, p, y0, x0, h, w, steps, crop_id=9000):
    """The crop is registered as an image of its own (its pixels and P4c) with
    the full-size sources; the oracle runs each half-step from the GPU's
    state and the winners / costs are compared as in test_gpu_pm."""
    img, P4c = _crop(site, ref, y0, x0, h, w)
    ctx = _ctx(site, p, srcs)
    try:
        ctx.add_image(crop_id, img, P4c)
        sids = [_id(site, s) for s in srcs]
        dr = site["dr"][ref]
        ov = _oracle(site, img, P4c, srcs, dr, crop_id, p)
        st = ctx.debug_init(crop_id, sids, dr, 0)
        rep = []
        yy, xx = np.mgrid[0:h, 0:w]
        for k in range(steps):
            it, color = k // 2, k % 2
            nxt = ctx.debug_halfstep(crop_id, sids, dr, st, 0, it, color)
            so = {key: st[key].astype(np.float64).copy() for key in ("depth", "normal", "cost", "prob")}
            _, cand = ov.halfstep(so, 0, it, color, diagnostics=True)
            act = ((xx + yy) % 2 == color).reshape(-1)
            dg, do = nxt["depth"][act], so["depth"][act]
            ng, no = nxt["normal"].reshape(-1, 3)[act], so["normal"].reshape(-1, 3)[act]
            same = (np.abs(dg - do) <= 1e-3 * (1 + np.abs(do))) & (np.abs(ng - no).max(1) <= 1e-3)
            srt = np.sort(cand[act], axis=1)
            decided = srt[:, 1] - srt[:, 0] > COST_TOL
            cerr = np.abs(nxt["cost"][act] - so["cost"][act])
            r = dict(step=k, decided_mismatch=float((~same & decided).mean()),
                     frac_cost_err=float((cerr > COST_TOL).mean()), cost_err_p999=float(np.quantile(cerr, 0.999)))
            print(r)
            rep.append(r)
            st = nxt
        return rep
    finally:
        ctx.close()


def _views(site, ids, k):
    from vissatsatellitestereo_amd.workspace import select_sources
    sel = select_sources(site["mvs"])
    return [dict(ref=_id(site, n), srcs=[_id(site, s) for s in sel[n][:k]], depth_range=site["dr"][n])
            for n in ids]


def _views_all(site, ids):
    """Every other view of the site as a source."""
    return [dict(ref=_id(site, n), srcs=[_id(site, s) for s in site["names"] if s != n],
                 depth_range=site["dr"][n]) for n in ids]


def _batched_equals_single(ctx, views, passes, want=("depth", "normal", "cost", "consistency")):
    from vissatsatellitestereo_amd.mvs import MvsContext  # noqa: F401
    for pass_ in passes:
        ctx.set_params(geom_consistency=pass_, filter=pass_)
        batch = ctx.solve(views, pass_, want=want, use_graph=True)
        for v, b in zip(views, batch):
            one = ctx.solve([v], pass_, want=want)[0]
            for k in want:
                assert np.array_equal(one[k], b[k]), (pass_, v["ref"], k)


# ---------------------------------------------------------------------------
# c2: 10 x 1024^2, 8 sources
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2(site_dir):
    from vissatsatellitestereo_amd import synth
    return sites.build("c2", synth.spec_c2(), site_dir, max_src=8)


def test_c2_cost_vs_oracle_full_image_coordinates(c2):
    ref = c2["names"][3]
    srcs = _sources(c2, ref, 8)
    assert len(srcs) == 8 and c2["imgs"][ref].shape == (1024, 1024)
    good = sites.gt_depth(c2, ref).astype(np.float32)
    for step in (1, 2):
        err, o = _cost_vs_oracle(c2, ref, srcs, _params(window_step=step), 600, 380, 128, 160, good)
        assert (o < 0.3).mean() > 0.15
        assert err.max() <= COST_TOL


@pytest.mark.parametrize("step", [1, 2])
def test_c2_lockstep_on_crop(c2, step):
    ref = c2["names"][0]
    srcs = _sources(c2, ref, 8)
    rep = _lockstep_crop(c2, ref, srcs, _params(window_step=step), 300, 420, 96, 128, 2)
    for r in rep:
        assert r["decided_mismatch"] <= 1e-3, r
        assert r["frac_cost_err"] <= 1e-3, r


def test_c2_batched_equals_single_full_size(c2):
    p = _params(window_step=2, num_iterations=1)
    ctx = _ctx(c2, p)
    try:
        _batched_equals_single(ctx, _views(c2, c2["names"], 8), (0, 1))
    finally:
        ctx.close()


def test_c2_dsm_error_against_ground_truth_and_oracle(c2):
    """Full 5 + 5 solve of the 10 views in both forms (batched at window step
    2, plain at 1), 0.5 m fusion: the median |DSM - ground truth| is within
    0.06 m and within 0.05 m of the FP64 oracle's on the same workload."""
    from vissatsatellitestereo_amd import runtime
    bar = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "c2_oracle_dsm.json")))
    grid = runtime.dsm_grid(runtime.read_aoi(c2["wd"]), 0.5, 0.5)
    gt = np.load(os.path.join(c2["wd"], "synthetic_gt", "gt_dsm.npy"))
    views = _views(c2, c2["names"], 8)
    ctx = _ctx(c2, _params())
    try:
        errs = {}
        for form, step in (("batched", 2), ("plain", 1)):
            ctx.set_params(window_step=step, geom_consistency=0, filter=0)
            if form == "batched":
                ctx.solve(views, 0, want=(), use_graph=True)
            else:
                for v in views:
                    ctx.solve([v], 0, want=())
            ctx.set_params(geom_consistency=1, filter=1)
            if form == "batched":
                ctx.solve(views, 1, want=(), use_graph=True)
            else:
                for v in views:
                    ctx.solve([v], 1, want=())
            dsm = ctx.fuse_dsm(grid, [v["ref"] for v in views], depth_kind=1, min_measurements=3)
            ok = ~np.isnan(dsm) & ~np.isnan(gt)
            errs[form] = float(np.median(np.abs(dsm[ok] - gt[ok])))
            assert ok.mean() > 0.9
    finally:
        ctx.close()
    print("c2 DSM median |dh|", errs, "oracle", bar["dsm_median_abs_err_m"])
    for form, e in errs.items():
        assert e <= 0.06, (form, e)
        assert e <= bar["dsm_median_abs_err_m"] + 0.05, (form, e)


# ---------------------------------------------------------------------------
# c3: 32 ragged views, 6..12 sources
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c3(site_dir):
    from vissatsatellitestereo_amd import synth
    return sites.build("c3", synth.spec_ragged(32), site_dir, max_src=12)


def test_c3_graph_batch_equals_single(c3):
    p = _params(window_step=2, num_iterations=1)
    ctx = _ctx(c3, p)
    try:
        rng = np.random.default_rng(7)
        counts = [int(k) for k in rng.integers(6, 13, len(c3["names"]))]
        from vissatsatellitestereo_amd.workspace import select_sources
        sel = select_sources(c3["mvs"])
        views = [dict(ref=_id(c3, n), srcs=[_id(c3, s) for s in sel[n][:k]], depth_range=c3["dr"][n])
                 for n, k in zip(c3["names"], counts)]
        sizes = [c3["imgs"][n].shape for n in c3["names"]]
        assert len(views) == 32 and min(min(s) for s in sizes) >= 512 and max(max(s) for s in sizes) <= 2048
        assert min(len(v["srcs"]) for v in views) >= 6 and max(len(v["srcs"]) for v in views) > 8
        ctx.set_params(geom_consistency=0, filter=0)
        batch = ctx.solve(views, 0, want=("depth", "normal"), use_graph=True)
        for v, b in zip(views, batch):
            one = ctx.solve([v], 0, want=("depth", "normal"))[0]
            assert np.array_equal(one["depth"], b["depth"]), v["ref"]
            assert np.array_equal(one["normal"], b["normal"]), v["ref"]
    finally:
        ctx.close()


# ---------------------------------------------------------------------------
# c4: serving, crops up to 3072^2, up to 16 sources
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c4(site_dir):
    from vissatsatellitestereo_amd import synth
    rng = np.random.default_rng(50)
    sizes = [(3072, 3072)] + [(int(rng.integers(1024, 1537)), int(rng.integers(1024, 1537))) for _ in range(16)]
    spec = synth.SceneSpec(n_views=17, sizes=sizes, off_nadir=(5.0, 30.0),
                           azimuths=tuple(45.0 * k for k in range(8)), n_boxes=30, n_points=3000,
                           seed_terrain=10, seed_albedo=11, seed_noise=12, seed_boxes=13, seed_cameras=14)
    return sites.build("c4", spec, site_dir, max_src=16, save_gt_views=False)


def _serve_until(ctx, ticket, ref, want=("depth", "normal", "cost")):
    for _ in range(10000):
        ctx.serve_step()
        r = ctx.poll(ticket, ref, want)
        if r is not None:
            return r
    raise AssertionError("request did not finish")


def test_c4_served_alone_equals_mixed_under_queueing(c4):
    from vissatsatellitestereo_amd.serving import ServeLoop
    p = _params(window_step=2, num_iterations=1)
    big = c4["names"][0]
    assert c4["imgs"][big].shape == (3072, 3072)
    views = _views(c4, [big, c4["names"][5], c4["names"][9]], 16)
    views[1]["srcs"] = views[1]["srcs"][:8]
    views[2]["srcs"] = views[2]["srcs"][:12]
    assert len(views[0]["srcs"]) >= 12
    ctx = _ctx(c4, p)
    try:
        ctx.set_params(filter=1)
        alone = [_serve_until(ctx, ctx.submit(v["ref"], v["srcs"], v["depth_range"]), v["ref"])
                 for v in views]
        # room for the 3072^2 request and the kept maps, not for a second request beside it
        h, w = c4["imgs"][big].shape
        base = ctx.stats()["hbm_bytes_in_use"]
        ctx.set_hbm_budget(base + h * w * (20 + 4 * 16 + 24) + 64 * 2**20)
        loop = ServeLoop(ctx, want=("depth", "normal", "cost"))
        t0 = loop.submit(views[1])
        loop.step()
        t1 = loop.submit(views[0])
        t2 = loop.submit(views[2])
        peak, queued = 0, False
        while loop.pending:
            na, nq = loop.step()
            peak = max(peak, na)
            queued |= nq > 0
        assert queued and peak < 3
        mixed = [loop.result(t1), loop.result(t0), loop.result(t2)]
        for i in range(3):
            for k in ("depth", "normal", "cost"):
                assert np.array_equal(alone[i][k], mixed[i][k]), (i, k)
    finally:
        ctx.close()


def test_c4_served_geometric_pass_equals_batched(c4):
    """Photometric requests for every view of the site, then geometric
    requests (filter on) for three of them including the 3072^2 crop with 16
    sources: served results equal the batched geometric solve."""
    p = _params(window_step=2, num_iterations=1)
    names = c4["names"]
    all_views = _views(c4, names, 16)
    geo = [all_views[0], all_views[4], all_views[11]]
    ref_ctx = _ctx(c4, p)
    try:
        ref_ctx.set_params(geom_consistency=0, filter=0)
        ref_ctx.solve(all_views, 0, want=(), use_graph=True)
        ref_ctx.set_params(geom_consistency=1, filter=1)
        batched = ref_ctx.solve(geo, 1, want=("depth", "normal", "consistency"), use_graph=True)
    finally:
        ref_ctx.close()
    ctx = _ctx(c4, p)
    try:
        ctx.set_params(geom_consistency=1)
        for v in all_views:
            ctx.submit(v["ref"], v["srcs"], v["depth_range"], pass_=0, filter=False)
        ts = [ctx.submit(v["ref"], v["srcs"], v["depth_range"], pass_=1, filter=True) for v in geo]
        for t, v, b in zip(ts, geo, batched):
            r = _serve_until(ctx, t, v["ref"], ("depth", "normal", "consistency"))
            for k in ("depth", "normal", "consistency"):
                assert np.array_equal(r[k], b[k]), (v["ref"], k)
            assert (r["depth"] > 0).mean() > 0.05
    finally:
        ctx.close()


# ---------------------------------------------------------------------------
# c5: 4096^2 views, 20 sources
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c5(site_dir):
    """21 views of the c5 generator (4096^2, 500 buildings): one reference
    view and 20 sources."""
    from vissatsatellitestereo_amd import synth
    return sites.build("c5", synth.spec_c5(n_views=21), site_dir, max_src=20, save_gt_views=False)


def _c5_good_depth(c5, ref, srcs, p):
    """Well-matched hypotheses without ground-truth views: the GPU's own depth
    after one photometric iteration at window step 2."""
    ctx = _ctx(c5, dict(p, num_iterations=1, window_step=2), [ref] + srcs)
    try:
        v = dict(ref=_id(c5, ref), srcs=[_id(c5, s) for s in srcs], depth_range=c5["dr"][ref])
        return ctx.solve([v], 0, want=("depth",))[0]["depth"]
    finally:
        ctx.close()


def test_c5_cost_vs_oracle_20_sources_stride_1(c5):
    ref = c5["names"][0]
    srcs = [s for s in c5["names"] if s != ref]
    assert len(srcs) == 20 and c5["imgs"][ref].shape == (4096, 4096)
    p = _params(window_step=1)
    good = _c5_good_depth(c5, ref, srcs, p)
    err, o = _cost_vs_oracle(c5, ref, srcs, p, 2900, 1700, 64, 80, good, seed=9)
    assert (o < 0.3).mean() > 0.05
    assert err.max() <= COST_TOL


@pytest.mark.parametrize("draw_group", [0, 1])
def test_c5_lockstep_20_sources_stride_1(c5, draw_group):
    ref = c5["names"][0]
    srcs = [s for s in c5["names"] if s != ref]
    rep = _lockstep_crop(c5, ref, srcs, _params(window_step=1, draw_group=draw_group), 1200, 2400, 48, 64, 2)
    for r in rep:
        assert r["decided_mismatch"] <= 1e-3, r
        assert r["frac_cost_err"] <= 1e-3, r


def test_c5_batched_equals_single_full_grid(c5):
    p = _params(window_step=2, num_iterations=1)
    ctx = _ctx(c5, p)
    try:
        views = _views_all(c5, c5["names"][:2])
        assert all(len(v["srcs"]) == 20 for v in views)
        _batched_equals_single(ctx, views, (0,), want=("depth", "normal", "cost"))
    finally:
        ctx.close()


def test_c5_fusion_0p3m_bit_exact(c5):
    """Depth maps of three 4096^2 views (one photometric iteration), fused on
    a 0.3 m grid over a 450 m window of the site: every per-view DSM and the
    fused DSM equal the numpy oracle's bit for bit."""
    import fusion_oracle as fo
    from vissatsatellitestereo_amd import runtime
    p = _params(window_step=2, num_iterations=1)
    names = c5["names"][:3]
    ctx = _ctx(c5, p)
    try:
        views = _views_all(c5, names)
        res = ctx.solve(views, 0, want=("depth",), use_graph=True)
        aoi = runtime.read_aoi(c5["wd"])
        full = runtime.dsm_grid(aoi, 0.3, 0.3)
        assert full[4] > 6000 and full[5] > 6000
        grid = (full[0] + 900 * 0.3, full[1] - 2400 * 0.3, 0.3, 0.3, 1500, 1500)
        dsm, per_view = ctx.fuse_dsm(grid, [v["ref"] for v in views], depth_kind=0, min_measurements=3,
                                     per_view=True)
    finally:
        ctx.close()
    ov = [fo.view_dsm(r["depth"], np.linalg.inv(c5["P4"][n]), grid) for r, n in zip(res, names)]
    for k in range(3):
        assert np.array_equal(per_view[k], ov[k], equal_nan=True), k
    want = fo.fuse(np.stack(ov), 3)
    assert np.array_equal(dsm, want, equal_nan=True)
    assert np.isfinite(dsm).mean() > 0.5
