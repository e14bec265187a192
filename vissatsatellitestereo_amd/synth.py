"""Seeded synthetic satellite-stereo scenes.

A scene is a heightfield (Perlin terrain plus rotated box buildings) with a
Perlin albedo texture, seen by perspective cameras ~600 km away with
f ~ 1e6 px, 5-30 deg off-nadir.  Images are rendered by ray casting in FP64 and
quantised to uint8 with Gaussian sensor noise.  The generator writes the same
inputs the reference's MVS stage consumes:

* a COLMAP PERSPECTIVE text model (cameras/images/points3D.txt, read by
  ``reparam_depth.py:73`` through ``colmap/read_model.py:261-265``) whose
  3D points are ray-cast surface points with visibility-checked tracks;
* grayscale PNG images (the ``sfm_perspective/images`` the dense workspace
  links to, ``stereo_pipeline.py:397-400``);
* ``aoi.json`` with the keys ``stereo_pipeline.py:210-223`` writes, in a local
  ENU frame (easting/northing = ENU east/north, alt = ENU up);
* ground truth: the 0.5 m DSM (per-cell max of the true surface) and, per
  view, the true hit height and camera depth for depth-map error statistics.

Everything is deterministic in the seeds (terrain, albedo, noise, boxes,
cameras).  Rendering runs on torch (CPU or GPU); it is data preparation, not
part of the measured path.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

# ----------------------------------------------------------------------------
# Perlin gradient noise (Perlin 2002, "Improving Noise"), vectorised.
# ----------------------------------------------------------------------------
_GRAD2 = np.array([[1, 1], [-1, 1], [1, -1], [-1, -1],
                   [1, 0], [-1, 0], [0, 1], [0, -1]], dtype=np.float64)


def _perm_table(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = rng.permutation(256).astype(np.int64)
    return np.concatenate([p, p])


def _fade(t):
    return t * t * t * (t * (t * 6.0 - 15.0) + 10.0)


def perlin2(x: np.ndarray, y: np.ndarray, perm: np.ndarray) -> np.ndarray:
    xi = np.floor(x).astype(np.int64)
    yi = np.floor(y).astype(np.int64)
    xf = x - xi
    yf = y - yi
    xi &= 255
    yi &= 255
    u = _fade(xf)
    v = _fade(yf)

    def g(ix, iy, dx, dy):
        h = perm[perm[ix] + iy] & 7
        gr = _GRAD2[h]
        return gr[..., 0] * dx + gr[..., 1] * dy

    n00 = g(xi, yi, xf, yf)
    n10 = g(xi + 1, yi, xf - 1.0, yf)
    n01 = g(xi, yi + 1, xf, yf - 1.0)
    n11 = g(xi + 1, yi + 1, xf - 1.0, yf - 1.0)
    nx0 = n00 + u * (n10 - n00)
    nx1 = n01 + u * (n11 - n01)
    return nx0 + v * (nx1 - nx0)


def fbm(x, y, seed, wavelength, octaves, persistence=0.5):
    out = np.zeros(np.broadcast(x, y).shape, dtype=np.float64)
    amp = 1.0
    lam = float(wavelength)
    for o in range(octaves):
        perm = _perm_table(seed * 7919 + o)
        out += amp * perlin2(x / lam + 0.37 * o, y / lam + 0.61 * o, perm)
        amp *= persistence
        lam *= 0.5
    return out


def fbm_grid(x0, y_top, res, n, seed, wavelength, octaves, persistence=0.5, device="cpu"):
    """fbm on the n x n raster with column 0 at easting x0 and row 0 at
    northing y_top (cell size res), computed with torch on `device`; the same
    values as fbm() on that raster up to FP64 rounding."""
    import torch
    dev = torch.device(device)
    xs = x0 + res * torch.arange(n, dtype=torch.float64, device=dev)
    ys = y_top - res * torch.arange(n, dtype=torch.float64, device=dev)
    Y, X = torch.meshgrid(ys, xs, indexing="ij")
    g2 = torch.tensor(_GRAD2, dtype=torch.float64, device=dev)
    out = torch.zeros((n, n), dtype=torch.float64, device=dev)
    amp = 1.0
    lam = float(wavelength)
    for o in range(octaves):
        perm = torch.from_numpy(_perm_table(seed * 7919 + o)).to(dev)
        x = X / lam + 0.37 * o
        y = Y / lam + 0.61 * o
        xi = torch.floor(x)
        yi = torch.floor(y)
        xf = x - xi
        yf = y - yi
        xi = xi.long() & 255
        yi = yi.long() & 255
        u = xf * xf * xf * (xf * (xf * 6.0 - 15.0) + 10.0)
        v = yf * yf * yf * (yf * (yf * 6.0 - 15.0) + 10.0)

        def gr(ix, iy, dx, dy):
            h = perm[perm[ix] + iy] & 7
            return g2[h, 0] * dx + g2[h, 1] * dy

        n00 = gr(xi, yi, xf, yf)
        n10 = gr(xi + 1, yi, xf - 1.0, yf)
        n01 = gr(xi, yi + 1, xf, yf - 1.0)
        n11 = gr(xi + 1, yi + 1, xf - 1.0, yf - 1.0)
        nx0 = n00 + u * (n10 - n00)
        nx1 = n01 + u * (n11 - n01)
        out += amp * (nx0 + v * (nx1 - nx0))
        amp *= persistence
        lam *= 0.5
    return out.cpu().numpy()


# ----------------------------------------------------------------------------
# Scene description
# ----------------------------------------------------------------------------
@dataclass
class SceneSpec:
    """Sizes and seeds of one synthetic site."""
    width: int = 256                 # image columns (all views unless sizes given)
    height: int = 256                # image rows
    n_views: int = 2
    off_nadir: tuple = (10.0, 20.0)  # degrees, uniform
    azimuths: tuple | None = None    # degrees; None -> evenly spread + jitter
    n_boxes: int = 20
    box_height: tuple = (5.0, 40.0)
    box_size: tuple = (8.0, 30.0)
    terrain_amp: float = 15.0
    noise_sigma: float = 2.0         # DN
    focal: float = 1.0e6             # px
    altitude: float = 6.0e5          # m
    aoi_fraction: float = 0.80       # AOI side / smallest image footprint side
    grid: float = 0.25               # heightfield raster spacing (m)
    sizes: list | None = None        # optional per-view (w, h)
    n_points: int = 1500             # sparse SfM points
    seed_terrain: int = 0
    seed_albedo: int = 1
    seed_noise: int = 2
    seed_boxes: int = 3
    seed_cameras: int = 4


@dataclass
class Camera:
    name: str
    width: int
    height: int
    K: np.ndarray      # 3x3
    R: np.ndarray      # 3x3 world->camera
    t: np.ndarray      # 3
    params: tuple = field(default=())   # fx, fy, cx, cy, s

    @property
    def center(self) -> np.ndarray:
        return -self.R.T @ self.t

    @property
    def P(self) -> np.ndarray:
        return self.K @ np.hstack([self.R, self.t.reshape(3, 1)])


def rotmat_to_qvec(R: np.ndarray) -> np.ndarray:
    """Unit quaternion (w, x, y, z) of a rotation matrix (Shepperd's method)."""
    tr = np.trace(R)
    if tr > 0:
        s = math.sqrt(tr + 1.0) * 2.0
        w = 0.25 * s
        x = (R[2, 1] - R[1, 2]) / s
        y = (R[0, 2] - R[2, 0]) / s
        z = (R[1, 0] - R[0, 1]) / s
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = math.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2.0
        w = (R[2, 1] - R[1, 2]) / s
        x = 0.25 * s
        y = (R[0, 1] + R[1, 0]) / s
        z = (R[0, 2] + R[2, 0]) / s
    elif R[1, 1] > R[2, 2]:
        s = math.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2.0
        w = (R[0, 2] - R[2, 0]) / s
        x = (R[0, 1] + R[1, 0]) / s
        y = 0.25 * s
        z = (R[1, 2] + R[2, 1]) / s
    else:
        s = math.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2.0
        w = (R[1, 0] - R[0, 1]) / s
        x = (R[0, 2] + R[2, 0]) / s
        y = (R[1, 2] + R[2, 1]) / s
        z = 0.25 * s
    q = np.array([w, x, y, z])
    q /= np.linalg.norm(q)
    if q[0] < 0:
        q = -q
    return q


def qvec_to_rotmat(q) -> np.ndarray:
    """Rotation matrix of a (w, x, y, z) quaternion, normalised first."""
    w, x, y, z = np.asarray(q, dtype=np.float64) / np.linalg.norm(q)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def make_camera(name, width, height, off_nadir_deg, azimuth_deg, focal,
                altitude, rng) -> Camera:
    th = math.radians(off_nadir_deg)
    ph = math.radians(azimuth_deg)
    to_cam = np.array([math.sin(th) * math.sin(ph), math.sin(th) * math.cos(ph),
                       math.cos(th)])
    C = (altitude / math.cos(th)) * to_cam
    zc = -to_cam
    e = np.array([1.0, 0.0, 0.0])
    xc = e - (e @ zc) * zc
    xc /= np.linalg.norm(xc)
    yc = np.cross(zc, xc)
    R = np.stack([xc, yc, zc])
    t = -R @ C
    fx = focal * (1.0 + rng.uniform(-0.01, 0.01))
    fy = focal * (1.0 + rng.uniform(-0.01, 0.01))
    s = rng.uniform(-20.0, 20.0)
    cx = (width - 1) / 2.0 + rng.uniform(-4.0, 4.0)
    cy = (height - 1) / 2.0 + rng.uniform(-4.0, 4.0)
    K = np.array([[fx, s, cx], [0.0, fy, cy], [0.0, 0.0, 1.0]])
    return Camera(name, width, height, K, R, t, (fx, fy, cx, cy, s))


class Scene:
    """Rasterised heightfield + albedo, and the cameras that view it."""

    def __init__(self, spec: SceneSpec, device="cpu"):
        self.spec = spec
        rng_cam = np.random.default_rng(spec.seed_cameras)
        sizes = spec.sizes or [(spec.width, spec.height)] * spec.n_views
        if spec.azimuths is None:
            base = np.linspace(0.0, 360.0, spec.n_views, endpoint=False)
            az = base + rng_cam.uniform(-15.0, 15.0, spec.n_views)
        else:
            az = np.resize(np.asarray(spec.azimuths, dtype=np.float64), spec.n_views)
            az = az + rng_cam.uniform(-5.0, 5.0, spec.n_views)
        off = rng_cam.uniform(spec.off_nadir[0], spec.off_nadir[1], spec.n_views)
        self.cameras = [make_camera(f"{i:04d}.png", int(sizes[i][0]), int(sizes[i][1]),
                                    off[i], az[i], spec.focal, spec.altitude, rng_cam)
                        for i in range(spec.n_views)]
        # Smallest ground footprint fixes the AOI (all views cover it).
        gsd = [np.linalg.norm(c.center) / c.K[0, 0] for c in self.cameras]
        side = min(min(c.width, c.height) * g for c, g in zip(self.cameras, gsd))
        L = math.floor(spec.aoi_fraction * side / 0.5) * 0.5
        self.aoi_side = L
        self.z_min = -spec.terrain_amp - 1.0
        self.z_max = spec.terrain_amp + spec.box_height[1] + 1.0
        # Raster domain: every image footprint at every height.
        ext = 0.0
        for c in self.cameras:
            for (u, v) in [(0, 0), (c.width - 1, 0), (0, c.height - 1), (c.width - 1, c.height - 1)]:
                for z in (self.z_min, self.z_max):
                    X = self._ray_at_height(c, u, v, z)
                    ext = max(ext, abs(X[0]), abs(X[1]))
        self.half = math.ceil(ext + 8.0)
        g = spec.grid
        n = int(round(2 * self.half / g)) + 1
        xs = -self.half + g * np.arange(n)
        self.grid_x0 = -self.half
        self.grid_n = n
        terr = fbm_grid(-self.half, -self.half + g * (n - 1), g, n, spec.seed_terrain,
                        wavelength=max(120.0, L * 0.8), octaves=5, persistence=0.5, device=device)
        terr *= spec.terrain_amp / max(1e-9, np.abs(terr).max())
        self.terrain = terr
        self.boxes = self._make_boxes(L)
        Hf = terr.copy()
        y_top = -self.half + g * (n - 1)
        for box in self.boxes:
            r0, r1, c0, c1 = self._box_window(box, xs[0], y_top, g, n, n)
            if r1 <= r0 or c1 <= c0:
                continue
            XX, YY = np.meshgrid(xs[c0:c1], y_top - g * np.arange(r0, r1), indexing="xy")
            m = self._in_box(XX, YY, *box[:5])
            blk = Hf[r0:r1, c0:c1]
            blk[m] = np.maximum(blk[m], box[5])
        self.heightfield = Hf
        # Albedo texture over a wider domain (walls sample it at x + 0.37 z, y + 0.61 z).
        self.alb_half = self.half + 45.0
        na = int(round(2 * self.alb_half / g)) + 1
        a = fbm_grid(-self.alb_half, -self.alb_half + g * (na - 1), g, na, spec.seed_albedo,
                     wavelength=24.0, octaves=6, persistence=0.65, device=device)
        a = a / max(1e-9, np.abs(a).max())
        self.albedo = 0.5 + 0.4 * a
        self.alb_n = na

    # ------------------------------------------------------------------
    @staticmethod
    def _ray_at_height(cam: Camera, u, v, z):
        d = cam.R.T @ np.linalg.solve(cam.K, np.array([u, v, 1.0]))
        C = cam.center
        s = (z - C[2]) / d[2]
        return C + s * d

    def _make_boxes(self, L):
        sp = self.spec
        rng = np.random.default_rng(sp.seed_boxes)
        boxes = []
        perm = None
        for _ in range(sp.n_boxes):
            cxb, cyb = rng.uniform(-0.45 * L, 0.45 * L, 2)
            w, l = rng.uniform(sp.box_size[0], sp.box_size[1], 2)
            ang = rng.uniform(0.0, math.pi / 2)
            h = rng.uniform(sp.box_height[0], sp.box_height[1])
            ti = self.terrain_at(np.array([cxb]), np.array([cyb]))[0]
            boxes.append((cxb, cyb, w, l, ang, ti + h))
        del perm
        return boxes

    @staticmethod
    def _box_window(box, x0, y_top, res, nrows, ncols):
        """Row/column range of a raster (row 0 at northing y_top, column 0 at
        easting x0, cell size res) that can contain the rotated box."""
        cxb, cyb, w, l = box[:4]
        rad = 0.5 * math.hypot(w, l) + res
        c0 = max(0, int(math.floor((cxb - rad - x0) / res)))
        c1 = min(ncols, int(math.ceil((cxb + rad - x0) / res)) + 1)
        r0 = max(0, int(math.floor((y_top - (cyb + rad)) / res)))
        r1 = min(nrows, int(math.ceil((y_top - (cyb - rad)) / res)) + 1)
        return r0, r1, c0, c1

    @staticmethod
    def _in_box(X, Y, cxb, cyb, w, l, ang):
        ca, sa = math.cos(ang), math.sin(ang)
        dx = X - cxb
        dy = Y - cyb
        a = ca * dx + sa * dy
        b = -sa * dx + ca * dy
        return (np.abs(a) <= w / 2) & (np.abs(b) <= l / 2)

    def terrain_at(self, x, y):
        g = self.spec.grid
        col = (x - self.grid_x0) / g
        row = (self.half - y) / g
        return _bilerp_np(self.terrain, row, col)

    def true_height(self, x, y):
        """Analytic surface height (terrain raster + exact box footprints)."""
        h = self.terrain_at(x, y)
        for (cxb, cyb, w, l, ang, roof) in self.boxes:
            rad = 0.5 * math.hypot(w, l)
            near = (np.abs(x - cxb) <= rad) & (np.abs(y - cyb) <= rad)
            if not near.any():
                continue
            m = near & self._in_box(x, y, cxb, cyb, w, l, ang)
            h = np.where(m, np.maximum(h, roof), h)
        return h

    # ------------------------------------------------------------------
    def aoi(self) -> dict:
        L = self.aoi_side
        return {"zone_number": 0, "hemisphere": "N",
                "ul_easting": -L / 2, "ul_northing": L / 2,
                "lr_easting": L / 2, "lr_northing": -L / 2,
                "width": L, "height": L,
                "lat_min": 0.0, "lat_max": 0.0, "lon_min": 0.0, "lon_max": 0.0,
                "alt_min": float(math.floor(self.z_min)), "alt_max": float(math.ceil(self.z_max)),
                "frame": "local ENU (east, north, up) metres"}

    def geo_aoi(self, lat_c=30.3, lon_c=-81.2) -> dict:
        """aoi.json in UTM (stereo_pipeline.py:185-226) for the site placed
        with its ENU origin near (lat_c, lon_c): the UTM box of side
        aoi_side is centred on the UTM coordinates of (lat_c, lon_c), the
        ENU origin is the centre of the box's lat/lon bounds
        (coordinate_system.py:41-51) at altitude alt_min = 0, so the
        scene's z is the altitude up to the Earth's curvature."""
        from . import geodesy
        L = self.aoi_side
        zone = geodesy.utm_zone(lat_c, lon_c)
        hemi = "N" if lat_c >= 0 else "S"
        E, N = geodesy.latlon_to_utm(lat_c, lon_c, zone, hemi == "N")
        return geodesy.aoi_dict(zone, hemi, float(np.round(E - L / 2, 1)), float(np.round(N + L / 2, 1)),
                                L, L, 0.0, float(math.ceil(self.z_max)))

    def gt_dsm_utm(self, aoi, resolution=0.5, sub=2) -> np.ndarray:
        """Per-cell max of the true surface altitude on the UTM grid of a geo
        aoi.json: each of sub x sub samples per cell goes UTM -> lat/lon ->
        the altitude where its geodetic vertical meets the surface (three
        fixed-point rounds of h += H(e, n) - u in the site's ENU frame)."""
        from . import geodesy
        lat0, lon0, alt0, zone, south = geodesy.geo_params_from_aoi(aoi)
        e_size = int(aoi["width"] / resolution) + 1
        n_size = int(aoi["height"] / resolution) + 1
        offs = (np.arange(sub) + 0.5) / sub
        out = np.full((n_size, e_size), -np.inf)
        cols, rows = np.arange(e_size), np.arange(n_size)
        for i in range(sub):
            for j in range(sub):
                EE, NN = np.meshgrid(aoi["ul_easting"] + (cols + offs[j]) * resolution,
                                     aoi["ul_northing"] - (rows + offs[i]) * resolution, indexing="xy")
                lat, lon = geodesy.utm_to_latlon(EE, NN, zone, south == 0)
                h = np.full(lat.shape, alt0)
                for _ in range(3):
                    e, n, u = geodesy.geodetic_to_enu(lat, lon, h, lat0, lon0, alt0)
                    h = h + (self.true_height(e, n) - u)
                out = np.maximum(out, h)
        return out

    def gt_dsm(self, resolution=0.5, sub=4) -> np.ndarray:
        """Per-cell max of the true surface on the AOI grid.

        Cell (row, col) covers northing (ul_n - (row+1) res, ul_n - row res]
        and easting [ul_e + col res, ul_e + (col+1) res), the cell that
        ``floor((ul_n - N)/res)``, ``floor((E - ul_e)/res)`` assigns
        (``lib/proj_to_grid.py:42-43``).  Terrain and each box are sampled at
        sub x sub points per cell; boxes only within their bounding windows."""
        aoi = self.aoi()
        e_size = int(aoi["width"] / resolution) + 1
        n_size = int(aoi["height"] / resolution) + 1
        offs = (np.arange(sub) + 0.5) / sub
        out = np.full((n_size, e_size), -np.inf)
        cols = np.arange(e_size)
        ul_e, ul_n = aoi["ul_easting"], aoi["ul_northing"]
        for i in range(sub):
            for j in range(sub):
                E = ul_e + (cols + offs[j]) * resolution
                for r0 in range(0, n_size, 512):
                    rows = np.arange(r0, min(n_size, r0 + 512))
                    N = ul_n - (rows + offs[i]) * resolution
                    EE, NN = np.meshgrid(E, N, indexing="xy")
                    out[rows] = np.maximum(out[rows], self.terrain_at(EE, NN))
        for box in self.boxes:
            r0, r1, c0, c1 = self._box_window(box, ul_e, ul_n, resolution, n_size, e_size)
            if r1 <= r0 or c1 <= c0:
                continue
            rr = np.arange(r0, r1)
            cc = np.arange(c0, c1)
            hit = np.zeros((r1 - r0, c1 - c0), dtype=bool)
            for i in range(sub):
                for j in range(sub):
                    EE, NN = np.meshgrid(ul_e + (cc + offs[j]) * resolution,
                                         ul_n - (rr + offs[i]) * resolution, indexing="xy")
                    hit |= self._in_box(EE, NN, *box[:5])
            blk = out[r0:r1, c0:c1]
            blk[hit] = np.maximum(blk[hit], box[5])
        return out

def _bilerp_np(img, row, col):
    h, w = img.shape
    row = np.clip(row, 0, h - 1.000001)
    col = np.clip(col, 0, w - 1.000001)
    r0 = np.floor(row).astype(np.int64)
    c0 = np.floor(col).astype(np.int64)
    fr = row - r0
    fc = col - c0
    a = img[r0, c0]
    b = img[r0, c0 + 1]
    c = img[r0 + 1, c0]
    d = img[r0 + 1, c0 + 1]
    return (1 - fr) * ((1 - fc) * a + fc * b) + fr * ((1 - fc) * c + fc * d)


# ----------------------------------------------------------------------------
# Ray-cast rendering (torch, FP64)
# ----------------------------------------------------------------------------
def _bilerp_t(img, row, col):
    import torch
    h, w = img.shape
    row = row.clamp(0, h - 1.000001)
    col = col.clamp(0, w - 1.000001)
    r0 = torch.floor(row)
    c0 = torch.floor(col)
    fr = row - r0
    fc = col - c0
    r0 = r0.long()
    c0 = c0.long()
    flat = img.reshape(-1)
    i00 = r0 * w + c0
    a = flat[i00]
    b = flat[i00 + 1]
    c = flat[i00 + w]
    d = flat[i00 + w + 1]
    return (1 - fr) * ((1 - fc) * a + fc * b) + fr * ((1 - fc) * c + fc * d)


def render_view(scene: Scene, cam: Camera, view_index: int, device="cpu"):
    """Ray-cast one view.  Returns (uint8 image, hit xyz float64 [H,W,3],
    camera depth float64 [H,W])."""
    import torch
    sp = scene.spec
    dev = torch.device(device)
    Ht = torch.from_numpy(scene.heightfield).to(dev)
    At = torch.from_numpy(scene.albedo).to(dev)
    g = sp.grid
    Kinv = np.linalg.inv(cam.K)
    M = cam.R.T @ Kinv
    C = cam.center
    vv, uu = torch.meshgrid(torch.arange(cam.height, dtype=torch.float64, device=dev),
                            torch.arange(cam.width, dtype=torch.float64, device=dev), indexing="ij")
    dx = M[0, 0] * uu + M[0, 1] * vv + M[0, 2]
    dy = M[1, 0] * uu + M[1, 1] * vv + M[1, 2]
    dz = M[2, 0] * uu + M[2, 1] * vv + M[2, 2]
    sx = dx / dz   # d x / d z along the ray
    sy = dy / dz

    def xy_at(z):
        return C[0] + (z - C[2]) * sx, C[1] + (z - C[2]) * sy

    def h_at(x, y):
        return _bilerp_t(Ht, (scene.half - y) / g, (x - scene.grid_x0) / g)

    slope = float(torch.sqrt(sx * sx + sy * sy).max())
    dzs = 0.5 * g / max(slope, 1e-3)
    z_top = scene.z_max + 0.5
    z_bot = scene.z_min - 0.5
    nsteps = int(math.ceil((z_top - z_bot) / dzs)) + 1
    lo = torch.full_like(uu, float("nan"))
    hi = torch.full_like(uu, float("nan"))
    zprev = z_top
    for k in range(1, nsteps + 1):
        z = max(z_top - k * dzs, z_bot)
        x, y = xy_at(z)
        below = (z <= h_at(x, y)) & torch.isnan(lo)
        lo = torch.where(below, torch.full_like(lo, z), lo)
        hi = torch.where(below, torch.full_like(hi, zprev), hi)
        zprev = z
        if z <= z_bot:
            break
    lo = torch.where(torch.isnan(lo), torch.full_like(lo, z_bot), lo)
    hi = torch.where(torch.isnan(hi), torch.full_like(hi, z_bot + 1e-3), hi)
    # Bisection on the bracket: lo at/below the surface, hi above it.
    for _ in range(30):
        mid = 0.5 * (lo + hi)
        x, y = xy_at(mid)
        under = mid <= h_at(x, y)
        lo = torch.where(under, mid, lo)
        hi = torch.where(under, hi, mid)
    z = 0.5 * (lo + hi)
    x, y = xy_at(z)
    # Shading: heightfield gradient normal, fixed sun, Lambertian.
    e = g
    hx = (h_at(x + e, y) - h_at(x - e, y)) / (2 * e)
    hy = (h_at(x, y + e) - h_at(x, y - e)) / (2 * e)
    nrm = torch.sqrt(hx * hx + hy * hy + 1.0)
    sun = np.array([0.35, 0.25, 0.90])
    sun = sun / np.linalg.norm(sun)
    lam = ((-hx * sun[0] - hy * sun[1] + sun[2]) / nrm).clamp(min=0.0)
    alb = _bilerp_t(At, (scene.alb_half - (y + 0.61 * z)) / g,
                    (x + 0.37 * z + scene.alb_half) / g)
    inten = alb * (0.35 + 0.65 * lam)
    gen = torch.Generator(device="cpu").manual_seed(sp.seed_noise * 100003 + view_index)
    noise = torch.randn(inten.shape, generator=gen, dtype=torch.float64).to(dev)
    dn = torch.clamp(torch.round(inten * 255.0 + sp.noise_sigma * noise), 0, 255)
    img = dn.to(torch.uint8).cpu().numpy()
    xyz = torch.stack([x, y, z], dim=-1).cpu().numpy()
    Xc = cam.R @ np.moveaxis(xyz, -1, 0).reshape(3, -1) + cam.t.reshape(3, 1)
    zcam = Xc[2].reshape(cam.height, cam.width)
    return img, xyz, zcam


# ----------------------------------------------------------------------------
# COLMAP text model + workspace export
# ----------------------------------------------------------------------------
def _project(cam: Camera, X: np.ndarray):
    x = cam.P @ np.vstack([X.T, np.ones((1, X.shape[0]))])
    return x[0] / x[2], x[1] / x[2], x[2]


def sparse_points(scene: Scene, hits: list, rng) -> tuple:
    """Surface points sampled in the smallest view (every other view covers
    its footprint) with visibility-checked tracks."""
    cams = scene.cameras
    n = scene.spec.n_points
    v0 = int(np.argmin([c.width * c.height for c in cams]))
    c0 = cams[v0]
    rows = rng.integers(0, c0.height, n)
    cols = rng.integers(0, c0.width, n)
    X = hits[v0][rows, cols]
    tracks = []
    for i in range(n):
        tracks.append([(v0, float(cols[i]), float(rows[i]))])
    for j, cam in enumerate(cams):
        if j == v0:
            continue
        u, v, w = _project(cam, X)
        ui = np.round(u).astype(np.int64)
        vi = np.round(v).astype(np.int64)
        ok = (ui >= 0) & (vi >= 0) & (ui < cam.width) & (vi < cam.height) & (w > 0)
        for i in np.nonzero(ok)[0]:
            if np.linalg.norm(hits[j][vi[i], ui[i]] - X[i]) < 0.75:
                tracks[i].append((j, float(u[i]), float(v[i])))
    keep = [i for i in range(n) if len(tracks[i]) >= 2]
    return X[keep], [tracks[i] for i in keep]


def _f(v):
    return repr(float(v))


def write_colmap_text(sparse_dir, cameras, X, tracks):
    os.makedirs(sparse_dir, exist_ok=True)
    with open(os.path.join(sparse_dir, "cameras.txt"), "w") as fp:
        fp.write("# Camera list with one line of data per camera:\n")
        fp.write("#   CAMERA_ID, MODEL, WIDTH, HEIGHT, PARAMS[]\n")
        for i, c in enumerate(cameras):
            fx, fy, cx, cy, s = c.params
            fp.write(f"{i + 1} PERSPECTIVE {c.width} {c.height} {_f(fx)} {_f(fy)} {_f(cx)} {_f(cy)} {_f(s)}\n")
    obs = [[] for _ in cameras]       # per image: (x, y, point3D_id)
    point_track = []
    for pid, tr in enumerate(tracks, start=1):
        t = []
        for (j, u, v) in tr:
            obs[j].append((u, v, pid))
            t.append((j + 1, len(obs[j]) - 1))
        point_track.append(t)
    with open(os.path.join(sparse_dir, "images.txt"), "w") as fp:
        fp.write("# Image list with two lines of data per image:\n")
        fp.write("#   IMAGE_ID, QW, QX, QY, QZ, TX, TY, TZ, CAMERA_ID, NAME\n")
        fp.write("#   POINTS2D[] as (X, Y, POINT3D_ID)\n")
        for i, c in enumerate(cameras):
            q = rotmat_to_qvec(c.R)
            fp.write(f"{i + 1} {_f(q[0])} {_f(q[1])} {_f(q[2])} {_f(q[3])} "
                     f"{_f(c.t[0])} {_f(c.t[1])} {_f(c.t[2])} {i + 1} {c.name}\n")
            fp.write(" ".join(f"{_f(u)} {_f(v)} {pid}" for (u, v, pid) in obs[i]) + "\n")
    with open(os.path.join(sparse_dir, "points3D.txt"), "w") as fp:
        fp.write("# 3D point list with one line of data per point:\n")
        fp.write("#   POINT3D_ID, X, Y, Z, R, G, B, ERROR, TRACK[] as (IMAGE_ID, POINT2D_IDX)\n")
        for pid, (x, t) in enumerate(zip(X, point_track), start=1):
            tr = " ".join(f"{a} {b}" for (a, b) in t)
            fp.write(f"{pid} {_f(x[0])} {_f(x[1])} {_f(x[2])} 128 128 128 0.5 {tr}\n")


def cameras_with_exact_pose(cameras):
    """Cameras whose R is the one the written quaternion encodes, so the
    ground truth matches what readers of the text model reconstruct."""
    out = []
    for c in cameras:
        q = rotmat_to_qvec(c.R)
        R = qvec_to_rotmat(q)
        out.append(Camera(c.name, c.width, c.height, c.K, R, c.t, c.params))
    return out


def build_site(spec: SceneSpec, work_dir: str, device="cpu", save_gt_views=True,
               log=None, geo=None, dsm_resolution=0.5) -> dict:
    """Render a site and lay it out as the reference pipeline's inputs.

    Layout written (``stereo_pipeline.py:396-405`` links ``colmap/mvs`` to it):
      work_dir/aoi.json
      work_dir/colmap/sfm_perspective/images/<name>.png
      work_dir/colmap/sfm_perspective/tri_ba/{cameras,images,points3D}.txt
      work_dir/synthetic_gt/{gt_dsm.npy, <name>.hit_z.npy, <name>.zcam.npy}
    geo=None writes a local-ENU aoi.json (zone 0); geo=(lat, lon) a UTM one
    (Scene.geo_aoi) and also synthetic_gt/gt_dsm_utm.npy on its grid.  The
    ground-truth DSMs have cells of dsm_resolution metres.
    """
    from PIL import Image
    scene = Scene(spec, device=device)
    scene.cameras = cameras_with_exact_pose(scene.cameras)
    img_dir = os.path.join(work_dir, "colmap", "sfm_perspective", "images")
    sparse_dir = os.path.join(work_dir, "colmap", "sfm_perspective", "tri_ba")
    gt_dir = os.path.join(work_dir, "synthetic_gt")
    for d in (img_dir, sparse_dir, gt_dir):
        os.makedirs(d, exist_ok=True)
    hits = []
    for i, cam in enumerate(scene.cameras):
        img, xyz, zcam = render_view(scene, cam, i, device=device)
        Image.fromarray(img, mode="L").save(os.path.join(img_dir, cam.name))
        hits.append(xyz)
        if save_gt_views:
            np.save(os.path.join(gt_dir, cam.name[:-4] + ".hit_z.npy"), xyz[..., 2].astype(np.float32))
            np.save(os.path.join(gt_dir, cam.name[:-4] + ".zcam.npy"), zcam.astype(np.float64))
        if log:
            log(f"rendered {cam.name} {cam.width}x{cam.height}")
    rng = np.random.default_rng(spec.seed_cameras + 17)
    X, tracks = sparse_points(scene, hits, rng)
    write_colmap_text(sparse_dir, scene.cameras, X, tracks)
    aoi = scene.aoi() if geo is None else scene.geo_aoi(*geo)
    with open(os.path.join(work_dir, "aoi.json"), "w") as fp:
        json.dump(aoi, fp, indent=2)
    # sub-cell sampling: 4 x 4 samples per 0.5 m cell, 2 x 2 for finer grids
    np.save(os.path.join(gt_dir, "gt_dsm.npy"),
            scene.gt_dsm(dsm_resolution, sub=4 if dsm_resolution >= 0.5 else 2))
    if geo is not None:
        np.save(os.path.join(gt_dir, "gt_dsm_utm.npy"), scene.gt_dsm_utm(aoi, dsm_resolution))
    return {"scene": scene, "aoi": aoi, "n_points": len(X)}


# Named configurations (image sizes, views, sources) used by tests and bench.
def spec_c1() -> SceneSpec:
    """2 views of 256x256, 10-20 deg off-nadir, 20 boxes."""
    return SceneSpec(width=256, height=256, n_views=2, off_nadir=(10.0, 20.0),
                     azimuths=(0.0, 180.0), n_boxes=20)


def spec_c2(n_views=10, size=1024) -> SceneSpec:
    """10 views of 1024x1024, 5-30 deg off-nadir at 8 azimuths."""
    return SceneSpec(width=size, height=size, n_views=n_views, off_nadir=(5.0, 30.0),
                     azimuths=tuple(45.0 * k for k in range(8)), n_boxes=20,
                     n_points=3000)


def spec_c5(n_views=40, size=4096) -> SceneSpec:
    """Large site: 40 views of 4096x4096 over ~2 km^2 with 500 buildings."""
    return SceneSpec(width=size, height=size, n_views=n_views, off_nadir=(5.0, 30.0),
                     azimuths=tuple(45.0 * k for k in range(8)), n_boxes=500, n_points=6000,
                     terrain_amp=15.0)


def spec_ragged(n_views=32, lo=512, hi=2048, seed=5, n_boxes=40, seed_offset=0) -> SceneSpec:
    """n_views crops with width and height drawn independently and uniformly
    from [lo, hi] (seeded), 5-30 deg off-nadir at 8 azimuths."""
    rng = np.random.default_rng(seed)
    sizes = [(int(rng.integers(lo, hi + 1)), int(rng.integers(lo, hi + 1))) for _ in range(n_views)]
    return SceneSpec(n_views=n_views, sizes=sizes, off_nadir=(5.0, 30.0),
                     azimuths=tuple(45.0 * k for k in range(8)), n_boxes=n_boxes, n_points=3000,
                     seed_terrain=0 + seed_offset, seed_albedo=1 + seed_offset,
                     seed_noise=2 + seed_offset, seed_boxes=3 + seed_offset,
                     seed_cameras=4 + seed_offset)
