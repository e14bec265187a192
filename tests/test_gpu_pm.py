"""HIP PatchMatch against the FP64 oracle (oracle/pm_oracle.c) on the GPU.

* cost of fixed hypotheses: |GPU - oracle| <= 1e-4 per (pixel, source);
* random initialisation: identical Philox streams -> hypotheses within FP32
  rounding of the oracle's;
* lockstep half-steps: the oracle runs one half-step from the GPU's own
  state; the winning hypothesis must agree wherever the oracle's candidate
  costs are separated by more than 1e-4 (the documented tie tolerance), and
  every final cost must agree within 1e-4 (plus a tiny allowance for the
  rare Monte-Carlo draw that lands within FP32 rounding of a CDF boundary);
* full solves: depth accuracy against the synthetic ground truth, and GPU
  vs oracle agreement statistics.
"""
import numpy as np
import pytest

import sites

pytestmark = pytest.mark.gpu

COST_TOL = 1e-4


def _ctx(**params):
    from vissatsatellitestereo_amd.mvs import MvsContext
    return MvsContext(device=0, **params)


def _oracle(site, ref, srcs, params, src_depths=None):
    import pm_oracle as po
    return po.OracleView(site["imgs"][ref], [site["imgs"][s] for s in srcs], site["P4"][ref],
                         [site["P4"][s] for s in srcs], site["dr"][ref], view_id=_id(site, ref),
                         src_depths=src_depths, params=params)


def _id(site, name):
    return site["names"].index(name) + 1


def _params(**kw):
    p = dict(window_radius=5, window_step=1, num_iterations=5, num_samples=15)
    p.update(kw)
    return p


def _gpu_params(p):
    keys = ("window_radius", "window_step", "num_samples", "num_iterations", "sigma_color",
            "ncc_sigma", "min_triangulation_angle", "incident_angle_sigma",
            "geom_consistency_regularizer", "geom_consistency_max_cost", "filter_min_ncc",
            "filter_min_triangulation_angle", "filter_geom_consistency_max_cost",
            "filter_min_num_consistent", "perturbation", "normal_perturbation",
            "max_normal_angle", "perturb_decay", "n_far", "seed", "draw_group", "tile_width")
    import pm_oracle as po
    full = po.default_params(**p)
    return {k: full[k] for k in keys if k in full}


def _setup(site, params, ref_i=0, n_src=None):
    names = site["names"]
    ref = names[ref_i]
    srcs = [n for n in names if n != ref][:n_src]
    ctx = _ctx(**_gpu_params(params))
    for n in names:
        ctx.add_image(_id(site, n), site["imgs"][n], site["P4"][n])
    return ctx, ref, srcs


@pytest.fixture(scope="module")
def c1(site_dir):
    from vissatsatellitestereo_amd import synth
    return sites.build("c1", synth.spec_c1(), site_dir)


@pytest.fixture(scope="module")
def tiny(site_dir):
    return sites.build("tiny", sites.tiny_spec(), site_dir)


@pytest.fixture(scope="module")
def many(site_dir):
    """19 views of 96x80: 18 sources per reference, the kernel instantiation
    for more than 16 sources."""
    from vissatsatellitestereo_amd import synth
    spec = synth.SceneSpec(width=96, height=80, n_views=19, off_nadir=(5.0, 30.0), n_boxes=4,
                           n_points=800)
    return sites.build("many", spec, site_dir)


@pytest.fixture(scope="module")
def ragged(site_dir):
    """4 views whose sizes are not multiples of the 32 x 16 tile: partial
    tiles at the right and bottom edges."""
    from vissatsatellitestereo_amd import synth
    spec = synth.SceneSpec(n_views=4, off_nadir=(10.0, 25.0), azimuths=(0.0, 90.0, 180.0, 270.0),
                           sizes=[(150, 100), (136, 94), (121, 110), (144, 87)], n_boxes=5,
                           n_points=800)
    return sites.build("ragged", spec, site_dir)


def _random_hyp(site, ref, seed=7):
    h, w = site["imgs"][ref].shape
    rng = np.random.default_rng(seed)
    dmin, dmax = site["dr"][ref]
    d = rng.uniform(dmin, dmax, (h, w))
    n = rng.normal(size=(h, w, 3)) * [0.3, 0.3, 1.0]
    n[..., 2] = np.abs(n[..., 2]) + 0.3
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    return d.astype(np.float32), n.astype(np.float32)


def test_cost_fixed_hypotheses_c1(c1):
    p = _params()
    ctx, ref, srcs = _setup(c1, p)
    d, n = _random_hyp(c1, ref)
    # near-true depths too, so most windows are well-matched (NCC ~ 1)
    gt = sites.gt_depth(c1, ref).astype(np.float32)
    d[::2] = gt[::2]
    n[::2] = [0, 0, 1]
    g = ctx.eval_cost(_id(c1, ref), [_id(c1, s) for s in srcs], c1["dr"][ref], d, n, 0)
    o = _oracle(c1, ref, srcs, p).cost_map(d.astype(np.float64), n.astype(np.float64), 0)
    err = np.abs(g - o)
    print("cost |gpu-oracle| max", err.max(), "mean", err.mean(), "frac matched(<0.2)",
          (o < 0.2).mean())
    assert (o < 0.3).mean() > 0.2          # the test covers well-matched windows
    assert err.max() <= COST_TOL, np.unravel_index(err.argmax(), err.shape)


@pytest.mark.parametrize("radius,step", [(4, 2), (5, 3)])
def test_cost_fixed_hypotheses_multisource_strided(tiny, radius, step):
    """Strided windows; r = 5 at step 3 visits offsets -5, -2, 1, 4 (no centre tap)."""
    p = _params(window_radius=radius, window_step=step)
    ctx, ref, srcs = _setup(tiny, p, ref_i=1)
    d, n = _random_hyp(tiny, ref, seed=3)
    g = ctx.eval_cost(_id(tiny, ref), [_id(tiny, s) for s in srcs], tiny["dr"][ref], d, n, 0)
    o = _oracle(tiny, ref, srcs, p).cost_map(d.astype(np.float64), n.astype(np.float64), 0)
    err = np.abs(g - o)
    print(f"cost r{radius} step{step} max", err.max())
    assert err.max() <= COST_TOL


def test_cost_fixed_hypotheses_18_sources(many):
    p = _params(window_radius=5, window_step=2)
    ctx, ref, srcs = _setup(many, p)
    assert len(srcs) == 18
    d, n = _random_hyp(many, ref, seed=5)
    g = ctx.eval_cost(_id(many, ref), [_id(many, s) for s in srcs], many["dr"][ref], d, n, 0)
    o = _oracle(many, ref, srcs, p).cost_map(d.astype(np.float64), n.astype(np.float64), 0)
    err = np.abs(g - o)
    print("cost 18 sources max", err.max())
    assert err.max() <= COST_TOL


def test_lockstep_photometric_18_sources(many):
    """Half-steps of the kernel for up to 32 sources (no LDS tap-weight table)
    against the oracle."""
    rep, _ = _lockstep(many, _params(window_radius=5, window_step=2), 0, None, 0, 4)
    for r in rep:
        assert r["decided_mismatch"] <= 1e-3, r
        assert r["frac_cost_err"] <= 1e-3, r


@pytest.mark.parametrize("tile_width", [16, 32])
def test_lockstep_photometric_grouped_draws_18_sources(many, tile_width):
    """Source draws shared by 32 x 4-pixel groups (draw_group = 1): the same
    lockstep agreement with the oracle, which keys its draws the same way, in
    the kernel for more than 16 sources at both of its tile shapes."""
    rep, _ = _lockstep(many, _params(window_radius=5, window_step=2, draw_group=1, tile_width=tile_width),
                       0, None, 0, 4)
    for r in rep:
        assert r["decided_mismatch"] <= 1e-3, r
        assert r["frac_cost_err"] <= 1e-3, r


def test_lockstep_photometric_12_sources(many):
    """Half-steps of the kernel for 9..16 sources against the oracle."""
    rep, _ = _lockstep(many, _params(window_radius=5, window_step=2), 0, 12, 0, 4)
    for r in rep:
        assert r["decided_mismatch"] <= 1e-3, r
        assert r["frac_cost_err"] <= 1e-3, r


def test_lockstep_geometric_18_sources(many):
    names = many["names"]
    srcs = [n for n in names if n != names[0]]
    rep, _ = _lockstep(many, _params(window_radius=5, window_step=2), 0, None, 1, 2,
                       src_depths_gpu=_src_depths(many, srcs))
    for r in rep:
        assert r["decided_mismatch"] <= 1e-3, r
        assert r["frac_cost_err"] <= 1e-3, r


def test_lds_staged_costs_equal_cached(tiny):
    """The measurement kernel that reads window taps from LDS-staged boxes of
    the sources returns the same costs as its cached-gather twin and as
    vsm_eval_cost, for near-true and random planes, stride 1 and 2."""
    for step in (1, 2):
        p = _params(window_radius=5, window_step=step)
        ctx, ref, srcs = _setup(tiny, p)
        gt = sites.gt_depth(tiny, ref).astype(np.float32)
        d, n = _random_hyp(tiny, ref, seed=11)
        mixed = d.copy()
        mixed[1::2] = gt[1::2]
        n[1::2] = [0, 0, 1]
        fronto = np.zeros_like(n)
        fronto[..., 2] = 1.0
        for dd, nn, must_stage in ((gt, fronto, True), (mixed, n, False)):
            args = (_id(tiny, ref), [_id(tiny, s) for s in srcs], tiny["dr"][ref], dd, nn)
            a, _, fa = ctx.eval_cost_staging(*args, stage=0, repeats=1)
            b, _, fb = ctx.eval_cost_staging(*args, stage=1, repeats=1)
            g = ctx.eval_cost(*args, 0)
            print("step", step, "staged fraction", fb)
            assert fa == 0.0 and (fb > 0.5 or not must_stage)
            assert np.array_equal(a, b)
            assert np.abs(a - g).max() <= 1e-6


def test_init_matches_oracle(c1):
    p = _params()
    ctx, ref, srcs = _setup(c1, p)
    st = ctx.debug_init(_id(c1, ref), [_id(c1, s) for s in srcs], c1["dr"][ref])
    ov = _oracle(c1, ref, srcs, p)
    so = ov.new_state()
    ov.init(so, 0)
    assert np.allclose(st["depth"], so["depth"], rtol=1e-6, atol=1e-5)
    nerr = np.abs(st["normal"] - so["normal"])
    print("init normal max err", nerr.max())
    assert nerr.max() <= 1e-5
    assert np.all(st["cost"] == 2.0) and np.all(st["prob"] == 0.5)


def _lockstep(site, p, ref_i, n_src, pass_, steps, src_depths_gpu=None):
    """Run `steps` half-steps on the GPU; after each, run the oracle from the
    GPU's previous state and compare."""
    import pm_oracle as po
    ctx, ref, srcs = _setup(site, p, ref_i, n_src)
    rid, sids = _id(site, ref), [_id(site, s) for s in srcs]
    ov = _oracle(site, ref, srcs, p)
    if pass_ == 1:
        for s in srcs:
            ctx.set_depth_map(_id(site, s), src_depths_gpu[s], kind=0)
        ov.set_source_depths([src_depths_gpu[s] for s in srcs])
    st = ctx.debug_init(rid, sids, site["dr"][ref], pass_)
    H, W = site["imgs"][ref].shape
    report = []
    for k in range(steps):
        it, color = k // 2, k % 2
        nxt = ctx.debug_halfstep(rid, sids, site["dr"][ref], st, pass_, it, color)
        so = {key: st[key].astype(np.float64).copy() for key in ("depth", "normal", "cost", "prob")}
        win, cand = ov.halfstep(so, pass_, it, color, diagnostics=True)
        yy, xx = np.mgrid[0:H, 0:W]
        act = ((xx + yy) % 2 == color).reshape(-1)
        dg, do = nxt["depth"][act], so["depth"][act]
        ng = nxt["normal"].reshape(-1, 3)[act]
        no = so["normal"].reshape(-1, 3)[act]
        same = (np.abs(dg - do) <= 1e-3 * (1 + np.abs(do))) & (np.abs(ng - no).max(1) <= 1e-3)
        c = cand[act]
        srt = np.sort(c, axis=1)
        gap = srt[:, 1] - srt[:, 0]
        decided = gap > COST_TOL
        cerr = np.abs(nxt["cost"][act] - so["cost"][act])
        perr = np.abs(nxt["prob"].reshape(-1, H * W)[:, act] - so["prob"].reshape(-1, H * W)[:, act])
        r = dict(step=k, mismatch=float((~same).mean()), decided_mismatch=float((~same & decided).mean()),
                 cost_err_max=float(cerr.max()), cost_err_p999=float(np.quantile(cerr, 0.999)),
                 frac_cost_err=float((cerr > COST_TOL).mean()), prob_err_max=float(perr.max()))
        print(r)
        report.append(r)
        st = nxt
    return report, st


def test_lockstep_photometric_c1(c1):
    rep, _ = _lockstep(c1, _params(), 0, None, 0, 6)
    for r in rep:
        assert r["decided_mismatch"] <= 2e-4, r
        assert r["frac_cost_err"] <= 2e-4, r
        assert r["cost_err_p999"] <= COST_TOL, r


def test_lockstep_photometric_multisource(tiny):
    rep, _ = _lockstep(tiny, _params(window_radius=4), 0, 3, 0, 6)
    for r in rep:
        assert r["decided_mismatch"] <= 1e-3, r
        assert r["frac_cost_err"] <= 1e-3, r


@pytest.mark.parametrize("tile_width", [16, 32, 64])
def test_lockstep_photometric_strided(tiny, tile_width):
    """r = 5 at window step 2: the 36-tap window of the batched form (tap
    weights kept in LDS) against the oracle, half-step by half-step, for every
    tile shape of the half-step kernel (16 x 32, 32 x 16, 64 x 8)."""
    rep, _ = _lockstep(tiny, _params(window_radius=5, window_step=2, tile_width=tile_width), 0, 3, 0, 6)
    for r in rep:
        assert r["decided_mismatch"] <= 1e-3, r
        assert r["frac_cost_err"] <= 1e-3, r


def test_full_solve_accuracy_and_oracle_agreement(c1):
    p = _params()
    ctx, ref, srcs = _setup(c1, p)
    res = ctx.solve([dict(ref=_id(c1, ref), srcs=[_id(c1, s) for s in srcs],
                          depth_range=c1["dr"][ref])], pass_=0)[0]
    gt = sites.gt_depth(c1, ref)
    ov = _oracle(c1, ref, srcs, p)
    so = ov.solve(0, p["num_iterations"])
    do = so["depth"].reshape(gt.shape)
    m = np.zeros(gt.shape, bool)
    m[8:-8, 8:-8] = True
    eg = np.abs(res["depth"] - gt)[m]
    eo = np.abs(do - gt)[m]
    print("GPU median |dd|", np.median(eg), "<1m", (eg < 1).mean(), " oracle", np.median(eo), (eo < 1).mean())
    assert np.median(eg) <= np.median(eo) + 0.05
    assert abs((eg < 1).mean() - (eo < 1).mean()) < 0.03
    # the same hypothesis at most pixels (independent FP32/FP64 runs)
    agree = np.abs(res["depth"] - do)[m] < 0.05
    print("GPU/oracle depth agreement", agree.mean())
    assert agree.mean() > 0.6
    assert np.all(np.isfinite(res["depth"])) and np.all(np.isfinite(res["normal"]))
    nz = res["normal"][m]
    assert np.allclose(np.linalg.norm(nz, axis=-1), 1, atol=1e-4)
    assert np.all(nz[:, 2] <= 0)


@pytest.mark.parametrize("tile_width", [16, 32, 64])
def test_batched_equals_single(tiny, tile_width):
    p = _params(window_radius=4, num_iterations=3, tile_width=tile_width)
    ctx, _, _ = _setup(tiny, p)
    views = []
    for r in tiny["names"]:
        srcs = [n for n in tiny["names"] if n != r]
        views.append(dict(ref=_id(tiny, r), srcs=[_id(tiny, s) for s in srcs], depth_range=tiny["dr"][r]))
    batch = ctx.solve(views, 0)
    graph = ctx.solve(views, 0, use_graph=True)
    for v, b, g in zip(views, batch, graph):
        one = ctx.solve([v], 0)[0]
        assert np.array_equal(one["depth"], b["depth"])
        assert np.array_equal(one["normal"], b["normal"])
        assert np.array_equal(g["depth"], b["depth"])


def test_staged_batch_equals_solve(tiny):
    """vsm_batch_begin / _step / _end (one half-step per call, with and
    without the hipGraph, reused across calls) give exactly vsm_solve's
    results for both passes; a discarded batch leaves the context usable."""
    p = _params(window_radius=4, num_iterations=3)
    ctx, _, _ = _setup(tiny, p)
    names = tiny["names"]
    views = [dict(ref=_id(tiny, r), srcs=[_id(tiny, s) for s in names if s != r],
                  depth_range=tiny["dr"][r]) for r in names]
    want = ("depth", "normal", "cost", "consistency")
    ref0 = ctx.solve(views, 0, want=want, use_graph=True)
    ctx.set_params(geom_consistency=1, filter=1)
    ref1 = ctx.solve(views, 1, want=want, use_graph=True)
    ctx.set_params(geom_consistency=0, filter=0)
    junk = ctx.batch(views, 0, use_graph=True)
    junk.step(2)
    junk.discard()
    for graph in (True, False):
        for pass_, ref in ((0, ref0), (1, ref1)):
            ctx.set_params(geom_consistency=pass_, filter=pass_)
            b = ctx.batch(views, pass_, use_graph=graph)
            left = [b.step(1) for _ in range(6)]
            assert left == [5, 4, 3, 2, 1, 0]
            assert b.step(1) == 0
            res = b.end(want)
            for i in range(len(views)):
                for k in want:
                    assert np.array_equal(res[i][k], ref[i][k]), (graph, pass_, i, k)


def test_lockstep_partial_tiles(ragged):
    """Reference 150 x 100 (partial tiles on two edges), window step 2."""
    assert ragged["imgs"][ragged["names"][0]].shape == (100, 150)
    rep, _ = _lockstep(ragged, _params(window_radius=5, window_step=2), 0, None, 0, 4)
    for r in rep:
        assert r["decided_mismatch"] <= 1e-3, r
        assert r["frac_cost_err"] <= 1e-3, r


def test_batched_equals_single_partial_tiles(ragged):
    p = _params(window_radius=5, window_step=2, num_iterations=2)
    ctx, _, _ = _setup(ragged, p)
    views = []
    for r in ragged["names"]:
        srcs = [n for n in ragged["names"] if n != r]
        views.append(dict(ref=_id(ragged, r), srcs=[_id(ragged, s) for s in srcs],
                          depth_range=ragged["dr"][r]))
    batch = ctx.solve(views, 0, use_graph=True)
    for v, b in zip(views, batch):
        one = ctx.solve([v], 0)[0]
        assert np.array_equal(one["depth"], b["depth"])
        assert np.array_equal(one["normal"], b["normal"])


def test_geometric_pass_and_filter(tiny):
    p = _params(window_radius=4, num_iterations=3)
    ctx, _, _ = _setup(tiny, p)
    names = tiny["names"]
    views = [dict(ref=_id(tiny, r), srcs=[_id(tiny, s) for s in names if s != r],
                  depth_range=tiny["dr"][r]) for r in names]
    photo = ctx.solve(views, 0)
    ctx.set_params(geom_consistency=1, filter=1)
    geo = ctx.solve(views, 1)
    for n, ph, ge in zip(names, photo, geo):
        gt = sites.gt_depth(tiny, n)
        valid = ge["depth"] > 0
        ep = np.abs(ph["depth"] - gt)
        eg = np.abs(ge["depth"] - gt)[valid]
        print(n, "photo med", np.median(ep), "geo(valid) med", np.median(eg), "kept", valid.mean())
        assert valid.mean() > 0.3
        assert np.median(eg) <= np.median(ep) + 0.05
        pc = np.array([bin(int(m)).count("1") for m in ge["consistency"].ravel()]).reshape(valid.shape)
        assert np.all(pc[~valid] < 2) and np.all(pc[valid] >= 2)


def _src_depths(site, srcs, seed=5):
    rng = np.random.default_rng(seed)
    out = {}
    for s in srcs:
        d = sites.gt_depth(site, s).astype(np.float32)
        d = d + rng.normal(0, 0.2, d.shape).astype(np.float32)
        d[rng.random(d.shape) < 0.05] = 0.0
        out[s] = d
    return out


def test_lockstep_geometric_multisource(tiny):
    names = tiny["names"]
    srcs = [n for n in names if n != names[0]]
    p = _params(window_radius=4)
    rep, st = _lockstep(tiny, p, 0, None, 1, 4, src_depths_gpu=_src_depths(tiny, srcs))
    for r in rep:
        assert r["decided_mismatch"] <= 1e-3, r
        assert r["frac_cost_err"] <= 1e-3, r


def test_lockstep_geometric_strided(tiny):
    names = tiny["names"]
    srcs = [n for n in names if n != names[0]]
    p = _params(window_radius=5, window_step=2)
    rep, st = _lockstep(tiny, p, 0, None, 1, 4, src_depths_gpu=_src_depths(tiny, srcs))
    for r in rep:
        assert r["decided_mismatch"] <= 1e-3, r
        assert r["frac_cost_err"] <= 1e-3, r


def test_filter_matches_oracle(tiny):
    """Consistency masks agree except where a source sits on a threshold."""
    import ctypes
    import pm_oracle as po
    names = tiny["names"]
    ref, srcs = names[0], names[1:]
    p = _params(window_radius=4)
    sd = _src_depths(tiny, srcs)
    ctx, _, _ = _setup(tiny, p)
    for s in srcs:
        ctx.set_depth_map(_id(tiny, s), sd[s], kind=0)
    ov = _oracle(tiny, ref, srcs, p, src_depths=[sd[s] for s in srcs])
    st = ov.solve(1, 2, state=None)                      # oracle hypotheses to filter
    g = ctx.debug_filter(_id(tiny, ref), [_id(tiny, s) for s in srcs], tiny["dr"][ref],
                         dict(depth=st["depth"].astype(np.float32), normal=st["normal"].astype(np.float32)), 1)
    so = {k: st[k].copy() for k in ("depth", "normal")}
    mo = ov.filter(so, 1)
    diff = np.nonzero(g["mask"].astype(np.uint64) != mo)[0]
    v, pr = ov._view(), ov._p()
    W = ov.W
    for i in diff:
        y, x = divmod(int(i), W)
        marg = []
        for s in range(len(srcs)):
            c = po.lib().orc_photo_cost(ctypes.byref(v), ctypes.byref(pr), x, y, st["depth"][i],
                                        st["normal"][3 * i:3 * i + 3].ctypes.data_as(ctypes.POINTER(ctypes.c_double)), s)
            e = po.lib().orc_geo_error(ctypes.byref(v), ctypes.byref(pr), x, y, st["depth"][i], s)
            marg.append(min(abs(1 - c - pr.filter_min_ncc), abs(e - pr.filter_geom_max_cost)))
        assert min(marg) < 1e-3, (x, y, marg)
    print("filter mask mismatches", len(diff), "of", len(mo), "kept", np.mean(st["depth"] > 0))
    assert len(diff) <= 1e-3 * len(mo) + 2
    keep_g = g["depth"] > 0
    keep_o = so["depth"] > 0
    assert np.mean(keep_g != keep_o) <= 1e-3


def test_get_depth_map_round_trip(tiny):
    """vsm_get_depth_map returns what vsm_set_depth_map stored and the solve's
    own depth output; unknown kinds / missing maps raise."""
    from vissatsatellitestereo_amd._lib import VsmError
    from vissatsatellitestereo_amd.mvs import MvsContext
    ctx = MvsContext(device=0, window_radius=3, num_iterations=2)
    try:
        names = tiny["names"]
        for i, n in enumerate(names):
            ctx.add_image(700 + i, tiny["imgs"][n], tiny["P4"][n])
        with pytest.raises(VsmError):
            ctx.get_depth_map(700, kind=0)
        rng = np.random.default_rng(4)
        d = rng.uniform(1, 5, tiny["imgs"][names[0]].shape).astype(np.float32)
        ctx.set_depth_map(700, d, kind=1)
        assert np.array_equal(ctx.get_depth_map(700, kind=1), d)
        v = dict(ref=701, srcs=[700, 702], depth_range=tiny["dr"][names[1]])
        out = ctx.solve([v], 0, want=("depth",))
        assert np.array_equal(ctx.get_depth_map(701, kind=0), out[0]["depth"])
    finally:
        ctx.close()


def test_api_rejects_bad_input_and_stays_usable(tiny):
    """Invalid parameters and requests raise (nothing reaches a kernel) and
    leave the context as it was: a valid solve afterwards matches a fresh
    context's."""
    from vissatsatellitestereo_amd._lib import VsmError
    from vissatsatellitestereo_amd.mvs import MvsContext
    names = tiny["names"]
    ids = [800 + i for i in range(len(names))]
    v = dict(ref=ids[0], srcs=ids[1:], depth_range=tiny["dr"][names[0]])
    ctx = MvsContext(device=0, window_radius=4, num_iterations=2)
    try:
        for i, n in enumerate(names):
            ctx.add_image(ids[i], tiny["imgs"][n], tiny["P4"][n])
        for bad in (dict(window_step=5), dict(window_radius=0), dict(window_radius=99),
                    dict(num_samples=17), dict(num_samples=0), dict(n_far=0)):
            with pytest.raises(VsmError):
                ctx.set_params(**bad)
        assert ctx.params.window_radius == 4 and ctx.params.window_step == 1
        bad_views = [dict(v, ref=999), dict(v, srcs=[999]), dict(v, srcs=[ids[0]]),
                     dict(v, srcs=list(range(900, 933))), dict(v, depth_range=(5.0, 5.0)),
                     dict(v, depth_range=(float("nan"), 10.0))]
        for bv in bad_views:
            with pytest.raises(VsmError):
                ctx.solve([bv], 0, want=("depth",))
        with pytest.raises(VsmError):          # empty batch
            ctx.solve([], 0, want=("depth",))
        with pytest.raises(VsmError):          # fusion of no views
            ctx.fuse_dsm((0.0, 0.0, 0.5, 0.5, 8, 8), [])
        with pytest.raises(VsmError):
            ctx.add_image(900, tiny["imgs"][names[0]], np.full((4, 4), np.nan))
        for shape in ((2, 32769), (1, 64)):    # wider than 32768 px; a single row
            with pytest.raises(VsmError):
                ctx.add_image(901, np.zeros(shape, np.uint8), tiny["P4"][names[0]])
        with pytest.raises(VsmError):
            ctx.set_depth_map(ids[0], np.ones((3, 3), np.float32), kind=0)
        got = ctx.solve([v], 0, want=("depth",))[0]["depth"]
    finally:
        ctx.close()
    fresh = MvsContext(device=0, window_radius=4, num_iterations=2)
    try:
        for i, n in enumerate(names):
            fresh.add_image(ids[i], tiny["imgs"][n], tiny["P4"][n])
        ref = fresh.solve([v], 0, want=("depth",))[0]["depth"]
    finally:
        fresh.close()
    assert np.array_equal(got, ref)
    # a 160 x 128 image holds 20 KB of pixels and a 40 KB pair texture
    small = MvsContext(device=0, hbm_budget_bytes=1 << 15)
    try:
        with pytest.raises(VsmError):
            small.add_image(1, tiny["imgs"][names[0]], tiny["P4"][names[0]])
    finally:
        small.close()
