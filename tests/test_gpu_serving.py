"""Serving and batching on the GPU: a request's result is the same alone, in
a mix with staggered arrivals, under a tight HBM budget (queued admission),
and as part of a batched vsm_solve."""
import numpy as np
import pytest

import sites

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tiny(site_dir):
    return sites.build("tiny", sites.tiny_spec(), site_dir)


def _ctx(site, budget=200 * 10**9, **params):
    from vissatsatellitestereo_amd.mvs import MvsContext
    p = dict(window_radius=4, num_iterations=3)
    p.update(params)
    c = MvsContext(device=0, hbm_budget_bytes=budget, **p)
    for i, n in enumerate(site["names"]):
        c.add_image(i + 1, site["imgs"][n], site["P4"][n])
    return c


def _views(site):
    ids = {n: i + 1 for i, n in enumerate(site["names"])}
    return [dict(ref=ids[n], srcs=[ids[s] for s in site["names"] if s != n], depth_range=site["dr"][n])
            for n in site["names"]]


def test_alone_equals_mixed_equals_batched(tiny):
    from vissatsatellitestereo_amd.serving import ServeLoop
    views = _views(tiny)
    ctx = _ctx(tiny)
    solo = []
    for v in views:
        loop = ServeLoop(ctx, want=("depth", "normal", "cost"))
        t = loop.submit(v)
        loop.drain()
        solo.append(loop.result(t))
    loop = ServeLoop(ctx, want=("depth", "normal", "cost"))
    t0 = loop.submit(views[2])
    loop.step()
    loop.step()
    t1 = loop.submit(views[0])
    loop.step()
    t2 = loop.submit(views[3])
    t3 = loop.submit(views[1])
    loop.drain()
    mixed = {2: loop.result(t0), 0: loop.result(t1), 3: loop.result(t2), 1: loop.result(t3)}
    batched = ctx.solve(views, 0, want=("depth", "normal", "cost"))
    for i in range(len(views)):
        for k in ("depth", "normal", "cost"):
            assert np.array_equal(solo[i][k], mixed[i][k]), (i, k)
            assert np.array_equal(solo[i][k], batched[i][k]), (i, k)
    ctx.close()


def _footprint(h, w, n_src):
    """Device bytes one served photometric request holds while active, as
    the library counts them: state (hypothesis, cost, per-source
    probabilities), outputs (depth, normal, cost, mask) and the image's kept
    depth and normal maps; plus a margin for tile lists and block rounding."""
    return h * w * (16 + 4 + 4 * n_src) + h * w * 24 + h * w * 16 + 64 * 1024


def test_tight_budget_queues_and_completes(tiny):
    """A budget with room for fewer requests than are submitted: requests
    wait for memory in FIFO order and still return the same results."""
    from vissatsatellitestereo_amd.serving import ServeLoop
    views = _views(tiny)
    probe = _ctx(tiny)
    ref = probe.solve(views, 0, want=("depth",))
    probe.close()
    ctx = _ctx(tiny)
    base = ctx.stats()["hbm_bytes_in_use"]
    h, w = tiny["imgs"][tiny["names"][0]].shape
    fp = _footprint(h, w, len(views) - 1)
    # images + one request, and the kept maps of every view
    ctx.set_hbm_budget(base + 1024 * 1024 + int(1.4 * fp) + len(views) * h * w * 16)
    loop = ServeLoop(ctx)
    ts = [loop.submit(v) for v in views]
    peak_active, queued_while_active = 0, False
    while loop.pending:
        na, nq = loop.step()
        peak_active = max(peak_active, na)
        queued_while_active |= na > 0 and nq > 0
    print("peak active", peak_active)
    assert 1 <= peak_active < len(views) and queued_while_active
    for t, r in zip(ts, ref):
        assert np.array_equal(loop.result(t)["depth"], r["depth"])
    ctx.close()


def test_request_that_cannot_fit_fails_instead_of_waiting(tiny):
    """With nothing active, a ready request that does not fit the budget
    fails with ENOMEM (it would otherwise wait forever); later requests that
    fit still run."""
    from vissatsatellitestereo_amd._lib import VsmError
    views = _views(tiny)
    ctx = _ctx(tiny)
    try:
        base = ctx.stats()["hbm_bytes_in_use"]
        h, w = tiny["imgs"][tiny["names"][0]].shape
        t_big = ctx.submit(views[0]["ref"], views[0]["srcs"], views[0]["depth_range"])
        ctx.set_hbm_budget(base + 1024 * 1024 + h * w * 8)    # below any request's footprint
        ctx.serve_step()
        with pytest.raises(VsmError) as e:
            ctx.poll(t_big, views[0]["ref"])
        assert e.value.code == -12
        ctx.set_hbm_budget(0)
        t = ctx.submit(views[1]["ref"], views[1]["srcs"], views[1]["depth_range"])
        res = None
        for _ in range(100):
            ctx.serve_step()
            res = ctx.poll(t, views[1]["ref"])
            if res is not None:
                break
        one = ctx.solve([views[1]], 0, want=("depth",))[0]
        assert res is not None and np.array_equal(res["depth"], one["depth"])
    finally:
        ctx.close()


def test_retire_failure_affects_only_its_request(tiny):
    """A request whose retirement fails (its filter's tile list cannot be
    allocated once the budget is cut below what is in use) ends with that
    error; a request retiring in the same step without the filter, and the
    serve steps after it, are unaffected.  With 16-wide half-step tiles the
    filter (32-wide tiles) builds its own tile list at retirement, which is
    the allocation that fails here; with 32-wide tiles it reuses the
    request's and allocates nothing."""
    from vissatsatellitestereo_amd._lib import VsmError
    views = _views(tiny)
    ctx = _ctx(tiny, num_iterations=2, tile_width=16)
    try:
        alone = {}
        for v in views[:2]:
            t = ctx.submit(v["ref"], v["srcs"], v["depth_range"], filter=False)
            while True:
                ctx.serve_step()
                r = ctx.poll(t, v["ref"], ("depth", "normal"))
                if r is not None:
                    alone[v["ref"]] = r
                    break
        ta = ctx.submit(views[0]["ref"], views[0]["srcs"], views[0]["depth_range"], filter=True)
        tb = ctx.submit(views[1]["ref"], views[1]["srcs"], views[1]["depth_range"], filter=False)
        ctx.serve_step()                         # both admitted, first half-step
        ctx.set_hbm_budget(1)                    # every later allocation fails
        for _ in range(3):
            na, nq = ctx.serve_step()
        assert na == 0 and nq == 0
        with pytest.raises(VsmError) as e:
            ctx.poll(ta, views[0]["ref"])
        assert e.value.code == -12
        rb = ctx.poll(tb, views[1]["ref"], ("depth", "normal"))
        assert rb is not None
        assert np.array_equal(rb["depth"], alone[views[1]["ref"]]["depth"])
        assert np.array_equal(rb["normal"], alone[views[1]["ref"]]["normal"])
        ctx.set_hbm_budget(0)
        tc = ctx.submit(views[0]["ref"], views[0]["srcs"], views[0]["depth_range"], filter=False)
        rc = None
        for _ in range(100):
            ctx.serve_step()
            rc = ctx.poll(tc, views[0]["ref"], ("depth",))
            if rc is not None:
                break
        assert rc is not None and np.array_equal(rc["depth"], alone[views[0]["ref"]]["depth"])
    finally:
        ctx.close()


def test_served_geometric_pass_equals_batched(tiny):
    """Photometric requests (no filter) and the geometric-consistency
    requests with the filter, all submitted at once in a shuffled order: each
    geometric request starts after the photometric requests of its view and
    sources and returns exactly the batched geometric result (depth, normals,
    consistency mask); the served maps are the device maps fusion reads."""
    from vissatsatellitestereo_amd.serving import ServeLoop
    views = _views(tiny)
    ref_ctx = _ctx(tiny)
    ref_ctx.set_params(filter=0)
    ref_ctx.solve(views, 0, want=())
    ref_ctx.set_params(geom_consistency=1, filter=1)
    batched = ref_ctx.solve(views, 1, want=("depth", "normal", "consistency"))
    ref_ctx.close()
    ctx = _ctx(tiny)
    try:
        ctx.set_params(geom_consistency=1)
        loop = ServeLoop(ctx, want=("depth", "normal", "consistency"), geometric=True)
        order = [2, 0, 3, 1]
        ts = {}
        ts[order[0]] = loop.submit(views[order[0]])
        loop.step()
        ts[order[1]] = loop.submit(views[order[1]])
        loop.step()
        for i in order[2:]:
            ts[i] = loop.submit(views[i])
        loop.drain()
        for i, v in enumerate(views):
            r = loop.result(ts[i])
            for k in ("depth", "normal", "consistency"):
                assert np.array_equal(r[k], batched[i][k]), (i, k)
            assert np.array_equal(ctx.get_depth_map(v["ref"], kind=1), r["depth"])
        assert all((batched[i]["depth"] > 0).mean() > 0.3 for i in range(len(views)))
    finally:
        ctx.close()


def test_geometric_request_without_photometric_maps_fails(tiny):
    """A geometric request whose sources have no photometric maps and no
    earlier photometric request pending fails (no eternal wait)."""
    from vissatsatellitestereo_amd._lib import VsmError
    views = _views(tiny)
    ctx = _ctx(tiny)
    try:
        v = views[0]
        t = ctx.submit(v["ref"], v["srcs"], v["depth_range"], pass_=1, filter=True)
        ctx.serve_step()
        with pytest.raises(VsmError) as e:
            ctx.poll(t, v["ref"])
        assert e.value.code == -22
    finally:
        ctx.close()


def test_autotune_picks_and_caches(tiny, tmp_path):
    from vissatsatellitestereo_amd import autotune
    views = _views(tiny)
    ctx = _ctx(tiny)
    cache = str(tmp_path / "tune.json")
    st, rep = autotune.tune(ctx, views, iters=3, radius=4, cache_path=cache)
    print("autotune", rep)
    assert st in (1, 2) and not rep["cached"] and rep["calib_views"] == 4
    assert rep["tile_width"] in (16, 32, 64) and set(rep["tile_seconds"]) == {"16", "32", "64"}
    assert rep["kept_fraction"][str(st)] >= rep["kept_fraction"]["1"] - rep["max_drop"]
    st2, rep2 = autotune.tune(ctx, views, iters=3, radius=4, cache_path=cache)
    assert st2 == st and rep2["cached"]
    ctx.close()


def test_served_result_becomes_device_depth_and_fuses(tiny):
    """A finished request's depth is the image's photometric depth map on the
    device (what incremental per-site fusion reads), and the on_complete hook
    sees every request once."""
    from vissatsatellitestereo_amd.serving import ServeLoop
    views = _views(tiny)
    ctx = _ctx(tiny, filter=1)
    seen = []
    loop = ServeLoop(ctx, want=("depth",), on_complete=lambda t, v, r: seen.append((t, v["ref"])))
    tickets = [loop.submit(v) for v in views]
    loop.drain()
    assert sorted(t for t, _ in seen) == sorted(tickets)
    for t, v in zip(tickets, views):
        assert np.array_equal(ctx.get_depth_map(v["ref"], kind=0), loop.result(t)["depth"])
    import json, os
    import fusion_oracle as fo
    grid = fo.grid_from_aoi(json.load(open(os.path.join(tiny["wd"], "aoi.json"))))
    ids = [v["ref"] for v in views]
    dsm = ctx.fuse_dsm(grid, ids, depth_kind=0, min_measurements=3)
    views_o = [fo.view_dsm(loop.result(t)["depth"], np.linalg.inv(tiny["P4"][n]), grid)
               for t, n in zip(tickets, tiny["names"])]
    assert np.array_equal(dsm, fo.fuse(np.stack(views_o), 3), equal_nan=True)
    ctx.close()


@pytest.fixture(scope="module")
def many(site_dir):
    from vissatsatellitestereo_amd import synth
    spec = synth.SceneSpec(width=96, height=80, n_views=19, off_nadir=(5.0, 30.0), n_boxes=4,
                           n_points=800)
    return sites.build("many", spec, site_dir)


def test_served_with_18_sources_equals_batched(many):
    """The kernel for more than 16 sources (view-selection costs in the
    context's global scratch, tap weights in LDS at window step 2) gives the
    same results served with staggered arrivals as in one batched solve."""
    from vissatsatellitestereo_amd.serving import ServeLoop
    views = _views(many)[:3]
    assert all(len(v["srcs"]) == 18 for v in views)
    ctx = _ctx(many, window_radius=5, window_step=2, num_iterations=2)
    try:
        loop = ServeLoop(ctx, want=("depth", "normal", "cost"))
        t0 = loop.submit(views[1])
        loop.step()
        t1 = loop.submit(views[0])
        loop.step()
        t2 = loop.submit(views[2])
        loop.drain()
        served = [loop.result(t1), loop.result(t0), loop.result(t2)]
        batched = ctx.solve(views, 0, want=("depth", "normal", "cost"))
        for i in range(3):
            for k in ("depth", "normal", "cost"):
                assert np.array_equal(served[i][k], batched[i][k]), (i, k)
    finally:
        ctx.close()
