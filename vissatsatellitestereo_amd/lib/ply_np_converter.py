"""Drop-in for the reference's lib/ply_np_converter.py: binary little-endian
PLY point clouds (vertex element only).

``np2ply(vertex, out_ply, color=None, comments=None, text=False,
use_double=False)`` (lib/ply_np_converter.py:38-68) writes x, y, z (float32,
or float64 with use_double), optional red, green, blue (uint8) and, as an
extension, optional nx, ny, nz (float32) the way COLMAP's fused.ply carries
them; ``text=True`` writes an ASCII PLY with 4 decimals.  ``ply2np(in_ply)``
(lib/ply_np_converter.py:71-93) returns (xyz, color or None, comments or
None) for files written here (binary little-endian or ASCII).
"""
import numpy as np


def np2ply(vertex, out_ply, color=None, comments=None, text=False, use_double=False, normal=None):
    v = np.asarray(vertex)
    n = v.shape[0]
    ft = "f8" if use_double else "f4"
    fields = [("x", ft), ("y", ft), ("z", ft)]
    if normal is not None:
        fields += [("nx", "f4"), ("ny", "f4"), ("nz", "f4")]
    if color is not None:
        fields += [("red", "u1"), ("green", "u1"), ("blue", "u1")]
    rec = np.zeros(n, dtype=[(k, "<" + t if t != "u1" else t) for k, t in fields])
    rec["x"], rec["y"], rec["z"] = v[:, 0], v[:, 1], v[:, 2]
    if normal is not None:
        nn = np.asarray(normal, np.float32)
        rec["nx"], rec["ny"], rec["nz"] = nn[:, 0], nn[:, 1], nn[:, 2]
    if color is not None:
        c = np.asarray(color).astype(np.uint8)
        rec["red"], rec["green"], rec["blue"] = c[:, 0], c[:, 1], c[:, 2]
    names = {"f4": "float", "f8": "double", "u1": "uchar"}
    head = ["ply", "format " + ("ascii" if text else "binary_little_endian") + " 1.0"]
    head += ["comment " + c for c in (comments or [])]
    head += [f"element vertex {n}"] + [f"property {names[t]} {k}" for k, t in fields] + ["end_header"]
    with open(out_ply, "wb") as fp:
        fp.write(("\n".join(head) + "\n").encode())
        if text:
            fmt = " ".join("%i" if t == "u1" else "%.4f" for _, t in fields)
            for r in rec:
                fp.write((fmt % tuple(r) + "\n").encode())
        else:
            fp.write(rec.tobytes())


def ply2np(in_ply):
    with open(in_ply, "rb") as fp:
        raw = fp.read()
    end = raw.index(b"end_header\n") + len(b"end_header\n")
    lines = raw[:end].decode().splitlines()
    fmt = next(l.split()[1] for l in lines if l.startswith("format"))
    comments = [l[len("comment "):] for l in lines if l.startswith("comment ")] or None
    n = next(int(l.split()[2]) for l in lines if l.startswith("element vertex"))
    types = {"float": "<f4", "double": "<f8", "uchar": "u1", "int": "<i4", "uint8": "u1", "float32": "<f4",
             "float64": "<f8"}
    props = [(l.split()[2], types[l.split()[1]]) for l in lines if l.startswith("property")]
    if fmt == "ascii":
        vals = np.array([row.split() for row in raw[end:].decode().split("\n") if row.strip()][:n],
                        dtype=np.float64).reshape(n, len(props))
        rec = {k: vals[:, i] for i, (k, _) in enumerate(props)}
    else:
        arr = np.frombuffer(raw, dtype=np.dtype(props), count=n, offset=end)
        rec = {k: arr[k] for k, _ in props}
    xyz = np.stack([rec["x"], rec["y"], rec["z"]], axis=1)
    color = (np.stack([rec["red"], rec["green"], rec["blue"]], axis=1).astype(np.uint8)
             if "red" in rec else None)
    return xyz, color, comments
