"""Host-side tests (CPU only): reparam_depth, workspace files, array IO,
source selection, the synthetic generator, and the C-ABI library exports."""
import ctypes
import os
import re

import numpy as np
import pytest

import sites

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def tiny(site_dir):
    return sites.build("tiny", sites.tiny_spec(), site_dir)


def test_robust_depth_range_golden():
    """reparam_depth.py:40-65 by hand: 50 values 1..50 -> sorted[int(0.02*50)=1]=2,
    sorted[int(0.98*50)=49]=50 -> (2-10, 50+30); a single value v -> (v-10, v+30);
    empty -> (-1e20, -1e20)."""
    from vissatsatellitestereo_amd.reparam_depth import robust_depth_range
    out = robust_depth_range({"a": list(range(50, 0, -1)), "b": [7.0], "c": []})
    assert out["a"] == (2 - 10, 50 + 30)
    assert out["b"] == (7.0 - 10, 7.0 + 30)
    assert out["c"] == (-1e20, -1e20)


def test_reparam_depth_matches_loop_oracle(tiny):
    import reparam_oracle as ro
    from vissatsatellitestereo_amd.reparam_depth import read_name_values
    mvs = tiny["mvs"]
    o = ro.reparam_depth(os.path.join(mvs, "sparse"))
    raw = read_name_values(os.path.join(mvs, "raw_depth.txt"))
    rep = read_name_values(os.path.join(mvs, "reparam_depth.txt"))
    lr = read_name_values(os.path.join(mvs, "last_rows.txt"))
    dr = read_name_values(os.path.join(mvs, "depth_ranges.txt"))
    for n in tiny["names"]:
        assert np.allclose(raw[n], o["raw_depth"][n], rtol=1e-12)
        assert np.allclose(rep[n], o["reparam_depth"][n], rtol=1e-9, atol=1e-9)
        assert np.allclose(lr[n], o["last_rows"][n], rtol=1e-12)
        assert np.allclose(dr[n], o["depth_ranges"][n], rtol=1e-9, atol=1e-9)
    plane = open(os.path.join(mvs, "reference_plane.txt")).read().split()
    assert np.allclose([float(v) for v in plane], o["reference_plane"], rtol=1e-12)


def test_reparam_depth_is_scaled_height(tiny):
    """d = depth_min (z - z0) / z_cam: ground truth depth of the rendered
    scene lies inside the common depth range."""
    for n in tiny["names"]:
        gt = sites.gt_depth(tiny, n)
        lo, hi = tiny["dr"][n]
        assert lo < np.percentile(gt, 1) and np.percentile(gt, 99) < hi


def test_projection_round_trip(tiny):
    """A ground-truth hit point projects through P4 to its pixel and depth."""
    from vissatsatellitestereo_amd import synth
    n = tiny["names"][0]
    P4 = tiny["P4"][n]
    hz = np.load(os.path.join(tiny["wd"], "synthetic_gt", n[:-4] + ".hit_z.npy"))
    gt = sites.gt_depth(tiny, n)
    inv = np.linalg.inv(P4)
    for (r, c) in [(10, 20), (64, 80), (100, 150)]:
        X = inv @ np.array([c, r, 1.0, gt[r, c]])
        X = X[:3] / X[3]
        assert abs(X[2] - hz[r, c]) < 1e-3
        y = P4 @ np.append(X, 1.0)
        assert abs(y[0] / y[2] - c) < 1e-6 and abs(y[1] / y[2] - r) < 1e-6
    del synth


def test_array_io_round_trip(tmp_path):
    from vissatsatellitestereo_amd.colmap.read_dense import read_array, write_array
    a = np.random.default_rng(0).normal(size=(5, 7)).astype(np.float32)
    b = np.random.default_rng(1).normal(size=(5, 7, 3)).astype(np.float32)
    write_array(a, tmp_path / "a.bin")
    write_array(b, tmp_path / "b.bin")
    assert np.array_equal(read_array(tmp_path / "a.bin"), a)
    assert np.array_equal(read_array(tmp_path / "b.bin"), b)
    raw = open(tmp_path / "b.bin", "rb").read()
    assert raw.startswith(b"7&5&3&")
    # column-major (w, h, c): w varies fastest, so the second float is
    # pixel (row 0, col 1), channel 0, and float w*h is channel 1 of (0, 0)
    vals = np.frombuffer(raw[6:], dtype="<f4")
    assert vals[0] == b[0, 0, 0] and vals[1] == b[0, 1, 0] and vals[7] == b[1, 0, 0]
    assert vals[35] == b[0, 0, 1]
    import fusion_oracle as fo
    assert np.array_equal(fo.read_array(tmp_path / "b.bin"), b)


def test_workspace_layout(tiny):
    mvs = tiny["mvs"]
    for p in ("images", "sparse", "stereo/depth_maps", "stereo/normal_maps",
              "stereo/consistency_graphs", "stereo/patch-match.cfg", "stereo/fusion.cfg",
              "raw_depth.txt", "reparam_depth.txt", "last_rows.txt", "reference_plane.txt",
              "depth_ranges.txt"):
        assert os.path.exists(os.path.join(mvs, p)), p
    assert os.path.islink(os.path.join(mvs, "images"))
    cfg = open(os.path.join(mvs, "stereo", "patch-match.cfg")).read().split("\n")
    assert cfg[0] == tiny["names"][0] and cfg[1] == "__auto__, 20"


def test_patch_match_cfg_forms(tmp_path):
    from vissatsatellitestereo_amd.workspace import parse_patch_match_cfg
    p = tmp_path / "pm.cfg"
    p.write_text("a.png\n__auto__, 8\nb.png\n__all__\nc.png\na.png, b.png\n")
    assert parse_patch_match_cfg(p) == [("a.png", ("__auto__", 8)), ("b.png", ("__all__",)),
                                        ("c.png", ["a.png", "b.png"])]


def test_auto_source_selection(tiny):
    from vissatsatellitestereo_amd.workspace import select_sources
    s = select_sources(tiny["mvs"])
    for ref in tiny["names"]:
        assert ref not in s[ref]
        assert 1 <= len(s[ref]) <= 3


def _write_model(d, centers, points):
    """A COLMAP text model with identity rotations (t = -C) and the given 3D
    points, each with the set of image names that observe it."""
    names = sorted(centers)
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "cameras.txt"), "w") as fp:
        for i, _ in enumerate(names):
            fp.write(f"{i + 1} PERSPECTIVE 100 100 1000 1000 50 50 0\n")
    obs = {n: [] for n in names}
    tracks = []
    for pid, (X, seen) in enumerate(points, start=1):
        t = []
        for n in sorted(seen):
            obs[n].append(pid)
            t.append((names.index(n) + 1, len(obs[n]) - 1))
        tracks.append((X, t))
    with open(os.path.join(d, "images.txt"), "w") as fp:
        for i, n in enumerate(names):
            C = centers[n]
            fp.write(f"{i + 1} 1 0 0 0 {-C[0]} {-C[1]} {-C[2]} {i + 1} {n}\n")
            fp.write(" ".join(f"10 10 {pid}" for pid in obs[n]) + "\n")
    with open(os.path.join(d, "points3D.txt"), "w") as fp:
        for pid, (X, t) in enumerate(tracks, start=1):
            tr = " ".join(f"{a} {b}" for a, b in t)
            fp.write(f"{pid} {X[0]} {X[1]} {X[2]} 128 128 128 0.5 {tr}\n")


def test_auto_source_selection_hand_derived(tmp_path):
    """__auto__, N (stereo_pipeline.py:424) on a hand-built model: sources
    ranked by the number of sparse points shared with the reference, ties by
    name, among those whose median triangulation angle over the shared points
    reaches 5 degrees.  Cameras 100 m from the origin; points at the origin see
    the pair A-S under the angle between their directions.
      B  30 deg, 10 shared         C   1 deg, 20 shared (rejected: 1 < 5)
      D  20 deg,  5 shared         E  10 deg, 10 shared (ties B, after it)
      F   3 deg on 4 + ~55 deg on 2 shared: median of [3,3,3,3,55,55] = 3 -> rejected
      G   3 deg on 2 + ~55 deg on 4 shared: median 55 -> accepted, 6 shared
    so the ranking is B, E, G, D."""
    import math
    from vissatsatellitestereo_amd.workspace import select_sources

    def cam(theta_deg):
        t = math.radians(theta_deg)
        return np.array([100 * math.sin(t), 0.0, 100 * math.cos(t)])

    C = {"A.png": cam(0), "B.png": cam(30), "C.png": cam(1), "D.png": cam(20), "E.png": cam(10),
         "F.png": cam(3), "G.png": cam(-3)}
    origin = lambda k: np.array([0.01 * k, 0.0, 0.0])      # angle ~ the camera angle
    pts = []
    pts += [(origin(k), {"A.png", "B.png"}) for k in range(10)]
    pts += [(origin(k), {"A.png", "C.png"}) for k in range(20)]
    pts += [(origin(k), {"A.png", "D.png"}) for k in range(5)]
    pts += [(origin(k), {"A.png", "E.png"}) for k in range(10)]
    for s, n_near in (("F.png", 4), ("G.png", 2)):
        mid = (C["A.png"] + C[s]) / 2 + np.array([0.0, 5.0, 0.0])    # between the two cameras
        pts += [(origin(k), {"A.png", s}) for k in range(n_near)]
        pts += [(mid + [0.01 * k, 0, 0], {"A.png", s}) for k in range(6 - n_near)]
    # points seen by sources only do not count for A
    pts += [(origin(k), {"B.png", "D.png"}) for k in range(30)]
    mvs = tmp_path / "mvs"
    _write_model(str(mvs / "sparse"), C, pts)
    os.makedirs(mvs / "stereo")
    with open(mvs / "stereo" / "patch-match.cfg", "w") as fp:
        fp.write("A.png\n__auto__, 3\nB.png\n__auto__, 20\nD.png\nA.png, C.png\n")
    sel = select_sources(str(mvs))
    assert sel["A.png"] == ["B.png", "E.png", "G.png"]
    # B shares 10 points with A (30 deg) and 30 with D (10 deg apart): D first
    assert sel["B.png"] == ["D.png", "A.png"]
    assert sel["D.png"] == ["A.png", "C.png"]
    # the threshold is the reference's min_triangulation_angle: at 0.5 deg C qualifies too
    assert select_sources(str(mvs), min_triangulation_angle=0.5)["A.png"] == ["C.png", "B.png", "E.png"]


def test_synth_is_deterministic(tmp_path):
    from vissatsatellitestereo_amd import synth
    spec = synth.SceneSpec(width=48, height=40, n_views=2, n_boxes=2, n_points=100)
    a = synth.build_site(spec, str(tmp_path / "a"))
    b = synth.build_site(spec, str(tmp_path / "b"))
    from PIL import Image
    for n in ("0000.png", "0001.png"):
        ia = np.array(Image.open(tmp_path / "a/colmap/sfm_perspective/images" / n))
        ib = np.array(Image.open(tmp_path / "b/colmap/sfm_perspective/images" / n))
        assert np.array_equal(ia, ib)
    assert a["aoi"] == b["aoi"]


def test_header_symbols_exported():
    """The built library loads (no device call) and exports every function
    include/vissat_mvs.h declares, and the binding lists each of them."""
    from vissatsatellitestereo_amd import _lib
    hdr = open(os.path.join(ROOT, "include", "vissat_mvs.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(vsm_\w+)\s*\(", hdr, re.M))
    assert len(declared) >= 20
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert declared == bound, declared ^ bound


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of the C structs have the sizes and field offsets the C
    compiler gives for include/vissat_mvs.h."""
    import subprocess
    from vissatsatellitestereo_amd import _lib
    checks = [("vsm_params", _lib.Params), ("vsm_view_req", _lib.ViewReq),
              ("vsm_view_out", _lib.ViewOut), ("vsm_dsm_grid", _lib.DsmGrid), ("vsm_stats", _lib.Stats),
              ("vsm_geo_params", _lib.GeoParams), ("vsm_geo_frame", _lib.GeoFrame),
              ("vsm_fuse3d_params", _lib.Fuse3dParams)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "vissat_mvs.h"', "int main(void){"]
    for cname, py in checks:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    k = 0
    for cname, py in checks:
        assert vals[k] == ctypes.sizeof(py), cname
        k += 1
        for f, _ in py._fields_:
            assert vals[k] == getattr(py, f).offset, (cname, f)
            k += 1


def test_default_params_are_reference_flags():
    from vissatsatellitestereo_amd import _lib
    p = _lib.default_params()
    assert p.window_radius == 3 and p.num_samples == 15 and p.num_iterations == 12
    assert p.min_triangulation_angle == 5.0 and p.filter == 0 and p.geom_consistency == 0
    assert abs(p.geom_consistency_regularizer - 0.3) < 1e-7 and p.geom_consistency_max_cost == 3.0
    assert abs(p.filter_min_ncc - 0.1) < 1e-7 and abs(p.filter_min_triangulation_angle - 4.99) < 1e-6
    assert p.filter_geom_consistency_max_cost == 1.0 and p.filter_min_num_consistent == 2


def test_no_device_raises_loudly():
    """Without a HIP device the product path fails instead of falling back."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from vissatsatellitestereo_amd.mvs import MvsContext
    from vissatsatellitestereo_amd._lib import VsmError
    with pytest.raises(VsmError):
        MvsContext(device=0)


def test_consistency_graph_records_round_trip(tmp_path):
    """Bitmask -> COLMAP consistency-graph records -> back, against a
    per-pixel loop."""
    from vissatsatellitestereo_amd.colmap import consistency_graph as cg
    rng = np.random.default_rng(9)
    mask = rng.integers(0, 1 << 5, (7, 11)).astype(np.uint32)
    mask[rng.random(mask.shape) < 0.4] = 0
    src = [4, 0, 9, 2, 7]
    path = str(tmp_path / "g.bin")
    cg.write(path, mask, src)
    got = cg.read(path)
    exp = {}
    for r in range(7):
        for c in range(11):
            ids = [src[b] for b in range(5) if (int(mask[r, c]) >> b) & 1]
            if ids:
                exp[(r, c)] = ids
    assert got == exp
    assert list(got) == sorted(got)          # row-major record order
    raw = np.fromfile(path, dtype="<i4")
    assert raw.size == sum(3 + len(v) for v in exp.values())
    cg.write(path, np.zeros((3, 3), np.uint32), src)
    assert cg.read(path) == {}


ANCHOR_PROBE = r"""
#include <cstdio>
#include <cmath>
#include <cstdint>
#include "vsm_device.h"
// the gathers' byte offset: signed 24-bit operands, low 32 bits of the product
static int32_t mul24(int32_t a, int32_t b)
{
    const int64_t sa = (int64_t)(int32_t)((uint32_t)a << 8) >> 8;
    const int64_t sb = (int64_t)(int32_t)((uint32_t)b << 8) >> 8;
    return (int32_t)(uint32_t)(uint64_t)(sa * sb);
}
int main()
{
    const double big[] = {0.0, 1.0, -1.0, 3.5e6, -3.5e6, 8.4e6, 1.0e7, -1.0e7, 1.0e12, -1.0e12,
                          INFINITY, -INFINITY, NAN};
    const int sizes[][2] = {{2, 2}, {256, 256}, {4096, 4096}, {32768, 32768}, {32768, 2}};
    long bad = 0, n = 0;
    for (auto &sz : sizes)
        for (double qx : big)
            for (double qy : big)
                for (double qz : {1e-300, 1.0, 0.0, -1.0, 1e300}) {
                    const int W = sz[0], H = sz[1];
                    int ax, ay;
                    vsm::source_anchor(qx, qy, qz, W, H, ax, ay);
                    if (ax < 0 || ax >= W || ay < 0 || ay >= H) { ++bad; continue; }
                    // every tap clamped to the image box, relative to the anchor
                    const int xs[] = {-ax, 0, W - 1 - ax}, ys[] = {-ay, 0, H - 1 - ay};
                    // pair texture: W + 1 texels of 2 bytes per row
                    const int P = W + 1;
                    const int Pb = 2 * P, baseb = (int)(uint32_t)(2u * (uint32_t)(ay * P + ax));
                    for (int ix : xs)
                        for (int iy : ys) {
                            const uint32_t off = (uint32_t)(mul24(iy, Pb) + ((ix << 1) + baseb));
                            const int64_t want = 2ll * ((int64_t)(ay + iy) * P + ax + ix);
                            // the 4-byte word at the offset stays inside the texture
                            if (want + 4 > 2ll * P * H) ++bad;
                            ++n;
                            if ((int64_t)off != want) ++bad;
                        }
                }
    printf("%ld %ld\n", n, bad);
    return bad != 0;
}
"""


def test_source_anchor_keeps_gather_offsets_in_range(tmp_path):
    """The per-tile source anchor stays inside the image for projections far
    outside it, behind the camera, at infinity or NaN, so the 24-bit offset
    arithmetic of the window gathers gives the exact texel offset for every
    tap clamped to the image box, and the tap's 4-byte word stays inside the
    pair texture (2-byte texels, W + 1 per row; images up to 32768 px a side)."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    src = tmp_path / "anchor.cpp"
    src.write_text(ANCHOR_PROBE)
    exe = tmp_path / "anchor"
    csrc = os.path.join(ROOT, "vissatsatellitestereo_amd", "csrc")
    subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O1", "-I", csrc, "-I",
                    os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    n, bad = (int(v) for v in out.stdout.split())
    assert n > 10000 and bad == 0 and out.returncode == 0, out.stdout
