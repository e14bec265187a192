"""COLMAP dense-array files (depth, normal and cost maps).

Format (``colmap/read_dense.py:36-51``): an ASCII header ``"{w}&{h}&{c}&"``
followed by raw little-endian float32 in column-major (w, h, c) order;
reading returns an (h, w) or (h, w, c) array.  ``write_array`` is the inverse
the adapted COLMAP binary performs when it saves maps.
"""
import numpy as np


def read_array(path):
    with open(path, "rb") as fid:
        head = b""
        n_amp = 0
        while n_amp < 3:
            ch = fid.read(1)
            if not ch:
                raise ValueError(f"{path}: truncated header")
            head += ch
            if ch == b"&":
                n_amp += 1
        width, height, channels = (int(v) for v in head.decode("ascii").split("&")[:3])
        data = np.fromfile(fid, dtype="<f4")
    if data.size != width * height * channels:
        raise ValueError(f"{path}: expected {width * height * channels} floats, found {data.size}")
    arr = data.reshape((width, height, channels), order="F")
    return np.transpose(arr, (1, 0, 2)).squeeze()


def write_array(array, path):
    a = np.asarray(array, dtype="<f4")
    if a.ndim == 2:
        a = a[:, :, None]
    height, width, channels = a.shape
    with open(path, "wb") as fid:
        fid.write(f"{width}&{height}&{channels}&".encode("ascii"))
        np.transpose(a, (1, 0, 2)).reshape(-1, order="F").tofile(fid)
