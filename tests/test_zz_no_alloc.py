"""vsm_proj_to_grid_noalloc (the rasteriser that allocates nothing: its cell
keys live in a per-device static array) against vsm_proj_to_grid, through
the C ABI, on two calls in a row, the second on a smaller grid with fewer
points: outputs bit-identical, so the second call sees none of the first
call's keys."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _points(rng, n, grid):
    ul_e, ul_n, e_res, n_res, e_size, n_size = grid
    e = rng.uniform(ul_e - 3 * n_res, ul_e + (e_size + 3) * n_res, n)
    nn = rng.uniform(ul_n - (n_size + 3) * e_res, ul_n + 3 * e_res, n)
    h = rng.normal(20.0, 8.0, n)
    h[rng.random(n) < 0.01] = np.nan          # missing heights are skipped by the max
    # several points per cell, some cells empty (hole fill)
    return np.stack([e, nn, h], axis=1)


def test_noalloc_equals_allocating_entry_point():
    import torch
    from vissatsatellitestereo_amd import _lib
    from vissatsatellitestereo_amd.mvs import MvsContext
    lib = _lib.load()
    ctx = MvsContext(device=0)
    rng = np.random.default_rng(11)
    cases = [((100.0, 900.0, 0.5, 0.5, 1500, 1200), 600_000),     # 1.8 M cells
             ((-40.0, 75.0, 0.3, 0.3, 301, 217), 20_000)]          # smaller grid, fewer points
    try:
        for grid, n in cases:
            pts = _points(rng, n, grid)
            want = ctx.proj_to_grid(pts, grid)                  # allocating entry point
            d_pts = torch.from_numpy(pts).to("cuda")
            d_out = torch.full((grid[5], grid[4]), -1.0, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            g = _lib.DsmGrid(*grid)
            rc = lib.vsm_proj_to_grid_noalloc(ctx.h, ctypes.c_void_p(d_pts.data_ptr()), n,
                                              ctypes.byref(g), ctypes.c_void_p(d_out.data_ptr()), None)
            assert rc == 0, lib.vsm_last_error(ctx.h)
            ctx.synchronize()
            got = d_out.cpu().numpy()
            print("grid", grid[4], "x", grid[5], "points", n, "filled", np.isfinite(got).mean())
            assert np.array_equal(got, want, equal_nan=True)
            assert np.isfinite(got).mean() > 0.5 and np.isnan(want).any()
        # a grid beyond the static scratch is refused before anything runs
        big = _lib.DsmGrid(0.0, 0.0, 0.5, 0.5, 4096, 1025)
        d_pts = torch.zeros((1, 3), dtype=torch.float64, device="cuda")
        rc = lib.vsm_proj_to_grid_noalloc(ctx.h, ctypes.c_void_p(d_pts.data_ptr()), 1,
                                          ctypes.byref(big), ctypes.c_void_p(d_pts.data_ptr()), None)
        assert rc == -22
    finally:
        ctx.close()
