"""CPU restatement of the 3D fusion of aggregate_3d.py (numpy).

TEST INFRASTRUCTURE (the oracle): imported only by tests/ as the checker.

The reference runs `colmap stereo_fusion --input_type geometric
--StereoFusion.min_num_pixels 4 --StereoFusion.max_reproj_error 2
--StereoFusion.max_depth_error 1.0 --StereoFusion.max_normal_error 10`
(aggregate_3d.py:44-52); COLMAP is not part of the reference tree, so its
exact traversal is not reproducible here (parity unpinned against it).  This
restates the order-independent rule the library implements, term by term in
the same FP64 / FP32 operation order:

  pixel p of image i, depth d > 0 -> X = inv(P4_i) (col, row, 1, d) / w;
  image j != i is consistent when P4_j X lands (nearest pixel q, in bounds)
  on a depth d_j > 0 whose back-projection Y is within max_depth_error (m)
  of X, P4_i Y re-projects within max_reproj_error (px) of p, and the world
  normals satisfy n_i . n_j >= cos(max_normal_error);
  a consistent j < i means the cluster is emitted by image j; otherwise, with
  at least min_num_pixels agreeing pixels (p included), p yields the mean of
  X and the Y's (FP64, j ascending), the normalised sum of the normals (FP32)
  and the rounded mean grey value.
"""
import math

import numpy as np


def _backproject(Mi, col, row, d):
    h = [((Mi[k, 0] * col + Mi[k, 1] * row) + Mi[k, 2]) + Mi[k, 3] * d for k in range(4)]
    return h[0] / h[3], h[1] / h[3], h[2] / h[3]


def _project(P, X, Y, Z):
    return tuple(((P[k, 0] * X + P[k, 1] * Y) + P[k, 2] * Z) + P[k, 3] for k in range(3))


def _rotate(R, n):
    R = np.asarray(R, dtype=np.float32)
    return tuple((R[k, 0] * n[..., 0] + R[k, 1] * n[..., 1]) + R[k, 2] * n[..., 2] for k in range(3))


def fuse_points(depths, normals, grays, P4s, invs, rots, min_num_pixels=4, max_reproj_error=2.0,
                max_depth_error=1.0, max_normal_error_deg=10.0):
    """depths[i] H x W float32, normals[i] H x W x 3 float32 (camera frame),
    grays[i] H x W uint8, P4s/invs 4 x 4 FP64, rots 3 x 3 camera -> world.
    Returns (xyz N x 3 FP64, normal N x 3 float32, grey N uint8)."""
    n = len(depths)
    cos_max = np.float32(math.cos(float(np.float32(max_normal_error_deg)) * math.pi / 180.0))
    mde, mre = float(np.float32(max_depth_error)), float(np.float32(max_reproj_error))
    out_x, out_n, out_g = [], [], []
    for i in range(n):
        D = np.asarray(depths[i], dtype=np.float32)
        H, W = D.shape
        rr, cc = np.nonzero((D > 0) & np.isfinite(D))
        col, row = cc.astype(np.float64), rr.astype(np.float64)
        X, Y, Z = _backproject(np.asarray(invs[i]), col, row, D[rr, cc].astype(np.float64))
        ok = np.isfinite(X) & np.isfinite(Y) & np.isfinite(Z)
        rr, cc, col, row, X, Y, Z = rr[ok], cc[ok], col[ok], row[ok], X[ok], Y[ok], Z[ok]
        na, nb, nc = _rotate(rots[i], np.asarray(normals[i], np.float32)[rr, cc])
        count = np.ones(rr.size, np.int64)
        owned = np.ones(rr.size, bool)
        sx, sy, sz = X.copy(), Y.copy(), Z.copy()
        snx, sny, snz = na.copy(), nb.copy(), nc.copy()
        sg = np.asarray(grays[i])[rr, cc].astype(np.int64)
        for j in range(n):
            if j == i:
                continue
            Dj = np.asarray(depths[j], dtype=np.float32)
            Hj, Wj = Dj.shape
            u0, u1, u2 = _project(np.asarray(P4s[j]), X, Y, Z)
            with np.errstate(divide="ignore", invalid="ignore"):
                cx = np.floor(u0 / u2 + 0.5)
                cy = np.floor(u1 / u2 + 0.5)
            m = (u2 > 0) & (cx >= 0) & (cy >= 0) & (cx < Wj) & (cy < Hj)
            qy = np.where(m, cy, 0).astype(np.int64)
            qx = np.where(m, cx, 0).astype(np.int64)
            dj = Dj[qy, qx]
            m &= (dj > 0) & np.isfinite(dj)
            Xj, Yj, Zj = _backproject(np.asarray(invs[j]), cx, cy, dj.astype(np.float64))
            ex, ey, ez = Xj - X, Yj - Y, Zj - Z
            with np.errstate(invalid="ignore"):
                m &= np.sqrt((ex * ex + ey * ey) + ez * ez) <= mde
                v0, v1, v2 = _project(np.asarray(P4s[i]), Xj, Yj, Zj)
                m &= v2 > 0
                rx, ry = v0 / v2 - col, v1 / v2 - row
                m &= np.sqrt(rx * rx + ry * ry) <= mre
            ma, mb, mc = _rotate(rots[j], np.asarray(normals[j], np.float32)[qy, qx])
            m &= ((na * ma + nb * mb) + nc * mc) >= cos_max
            if j < i:
                owned &= ~m
                continue
            m &= owned
            count += m
            sx = np.where(m, sx + Xj, sx)
            sy = np.where(m, sy + Yj, sy)
            sz = np.where(m, sz + Zj, sz)
            snx = np.where(m, snx + ma, snx)
            sny = np.where(m, sny + mb, sny)
            snz = np.where(m, snz + mc, snz)
            sg = np.where(m, sg + np.asarray(grays[j])[qy, qx].astype(np.int64), sg)
        keep = owned & (count >= min_num_pixels)
        fc = count[keep].astype(np.float64)
        snx, sny, snz = snx[keep], sny[keep], snz[keep]
        l = np.sqrt((snx * snx + sny * sny) + snz * snz)
        with np.errstate(invalid="ignore", divide="ignore"):
            nrm = np.stack([np.where(l > 0, snx / l, np.float32(0)), np.where(l > 0, sny / l, np.float32(0)),
                            np.where(l > 0, snz / l, np.float32(0))], axis=1).astype(np.float32)
        out_x.append(np.stack([sx[keep] / fc, sy[keep] / fc, sz[keep] / fc], axis=1))
        out_n.append(nrm)
        cnt = count[keep]
        out_g.append(np.minimum((sg[keep] + cnt // 2) // cnt, 255).astype(np.uint8))
    return (np.concatenate(out_x) if out_x else np.zeros((0, 3)),
            np.concatenate(out_n) if out_n else np.zeros((0, 3), np.float32),
            np.concatenate(out_g) if out_g else np.zeros(0, np.uint8))


def camera_to_world(P4):
    """R^T of the camera rotation R in P4[:3, :3] = K R (K upper triangular,
    positive diagonal): the rotation taking camera-frame normals to the world."""
    A = np.asarray(P4, dtype=np.float64)[:3, :3]
    r3 = A[2] / np.linalg.norm(A[2])
    a2 = A[1] - (A[1] @ r3) * r3
    r2 = a2 / np.linalg.norm(a2)
    a1 = A[0] - (A[0] @ r3) * r3 - (A[0] @ r2) * r2
    r1 = a1 / np.linalg.norm(a1)
    return np.stack([r1, r2, r3]).T
