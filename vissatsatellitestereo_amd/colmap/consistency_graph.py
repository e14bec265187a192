"""Consistency graphs of the geometric pass in the dense workspace layout.

``stereo/consistency_graphs/<img>.geometric.bin`` (the directory the
reference creates at stereo_pipeline.py:415 for the external binary) holds,
as raw little-endian int32 without a header, one record per pixel that has
at least one consistent source: ``row, col, n, idx_1 .. idx_n`` in row-major
pixel order, idx being the 0-based position of the source image in the
workspace's image list (images sorted by name).  This is the public COLMAP
ConsistencyGraph layout that its stereo_fusion reads (COLMAP itself is not
part of the reference tree).

The device returns, per pixel, a bitmask over the view's source list (bit s
set: source s passed the filter's consistency test); ``write`` expands it.
"""
import numpy as np


def records(mask, src_indices):
    """int32 record stream of a H x W uint32 bitmask over the sources whose
    workspace indices are src_indices (bit s <-> src_indices[s])."""
    mask = np.asarray(mask, dtype=np.uint32)
    src = np.asarray(src_indices, dtype=np.int32)
    rows, cols = np.nonzero(mask)
    if rows.size == 0:
        return np.zeros(0, np.int32)
    m = mask[rows, cols]
    bits = ((m[:, None] >> np.arange(len(src), dtype=np.uint32)[None, :]) & 1).astype(bool)
    n = bits.sum(axis=1).astype(np.int64)
    keep = n > 0
    rows, cols, bits, n = rows[keep], cols[keep], bits[keep], n[keep]
    rec = 3 + n
    starts = np.concatenate([[0], np.cumsum(rec)[:-1]])
    out = np.empty(int(rec.sum()), np.int32)
    out[starts] = rows
    out[starts + 1] = cols
    out[starts + 2] = n
    pix, b = np.nonzero(bits)                      # row-major: pixel, then source order
    first = np.concatenate([[0], np.cumsum(n)[:-1]])
    rank = np.arange(pix.size) - first[pix]
    out[starts[pix] + 3 + rank] = src[b]
    return out


def write(path, mask, src_indices):
    records(mask, src_indices).astype("<i4").tofile(path)


def read(path):
    """{(row, col): [idx, ...]} from a consistency-graph file."""
    data = np.fromfile(path, dtype="<i4")
    out, k = {}, 0
    while k < data.size:
        r, c, n = int(data[k]), int(data[k + 1]), int(data[k + 2])
        out[(r, c)] = [int(v) for v in data[k + 3:k + 3 + n]]
        k += 3 + n
    return out
