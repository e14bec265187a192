"""Drop-in for the reference's lib/dsm_util.py: single-band float32 GeoTIFF DSMs.

``write_dsm_tif(image, out_file, geo, utm_zone, nodata_val=None)``
(lib/dsm_util.py:122-159) writes an uncompressed little-endian TIFF with the
GeoTIFF tags GDAL would write for that call: ModelPixelScale (e_res, n_res,
0), ModelTiepoint (0, 0, 0, ul_e, ul_n, 0), GeoKeys for a projected model,
pixel-is-area raster and EPSG 326zz / 327zz (WGS 84 / UTM zone zz N / S), the
GDAL metadata AREA_OR_POINT=Area and the GDAL nodata tag.  NaN cells are
written as nodata_val when one is given.  A zone number of 0 (a local ENU
grid) is written with the user-defined projected CS code 32767.

``read_dsm_tif(file)`` (lib/dsm_util.py:43-97) reads such a file back into
(image with nodata -> NaN, meta dict with the reference's keys).  Only
uncompressed single-band float32 strips are read.
"""
import os
import struct

import numpy as np

_T_ASCII, _T_SHORT, _T_LONG, _T_DOUBLE = 2, 3, 4, 12
_SIZES = {1: 1, 2: 1, 3: 2, 4: 4, 12: 8}


def _epsg(zone_number, hemisphere):
    if int(zone_number) == 0:
        return 32767
    return (32600 if hemisphere == "N" else 32700) + int(zone_number)


def parse_proj_str(proj_str):
    """(zone_number, hemisphere) from 'WGS 84 / UTM zone 17N'-like text
    (lib/dsm_util.py:32-40)."""
    idx = proj_str.find("UTM zone")
    sub = proj_str[idx + len("UTM zone"):].strip()
    digits = "".join(ch for ch in sub[:2] if ch.isdigit())
    return int(digits), sub[len(digits)]


def write_dsm_tif(image, out_file, geo, utm_zone, nodata_val=None):
    assert image.ndim == 2
    ul_e, ul_n, e_res, n_res = (float(v) for v in geo)
    zone_number, hemisphere = utm_zone
    img = np.ascontiguousarray(image, dtype=np.float32)
    if nodata_val is not None:
        img = img.copy()
        img[np.isnan(img)] = nodata_val
        nodata_txt = "%.17g" % float(nodata_val)
    else:
        nodata_txt = "nan"
    h, w = img.shape
    data = img.astype("<f4").tobytes()
    citation = ("WGS 84 / UTM zone {}{}|".format(zone_number, hemisphere) if int(zone_number)
                else "local ENU|").encode() + b"\0"
    keys = [1, 1, 0, 4,
            1024, 0, 1, 1,                          # GTModelTypeGeoKey: projected
            1025, 0, 1, 1,                          # GTRasterTypeGeoKey: pixel is area
            1026, 34737, len(citation) - 1, 0,      # GTCitationGeoKey -> ascii params
            3072, 0, 1, _epsg(zone_number, hemisphere)]
    meta = b'<GDALMetadata>\n  <Item name="AREA_OR_POINT">Area</Item>\n</GDALMetadata>\n\0'
    entries = [
        (256, _T_LONG, [w]), (257, _T_LONG, [h]), (258, _T_SHORT, [32]), (259, _T_SHORT, [1]),
        (262, _T_SHORT, [1]), (273, _T_LONG, [8]), (277, _T_SHORT, [1]), (278, _T_LONG, [h]),
        (279, _T_LONG, [len(data)]), (284, _T_SHORT, [1]), (339, _T_SHORT, [3]),
        (33550, _T_DOUBLE, [e_res, n_res, 0.0]),
        (33922, _T_DOUBLE, [0.0, 0.0, 0.0, ul_e, ul_n, 0.0]),
        (34735, _T_SHORT, keys),
        (34737, _T_ASCII, citation),
        (42112, _T_ASCII, meta),
        (42113, _T_ASCII, nodata_txt.encode() + b"\0"),
    ]
    ifd_off = 8 + len(data)
    ifd_off += ifd_off & 1
    extra_off = ifd_off + 2 + 12 * len(entries) + 4
    ifd = struct.pack("<H", len(entries))
    extra = b""
    for tag, typ, vals in entries:
        if typ == _T_ASCII:
            raw, count = bytes(vals), len(vals)
        else:
            fmt = {_T_SHORT: "H", _T_LONG: "I", _T_DOUBLE: "d"}[typ]
            raw, count = struct.pack("<%d%s" % (len(vals), fmt), *vals), len(vals)
        if len(raw) <= 4:
            ifd += struct.pack("<HHI", tag, typ, count) + raw.ljust(4, b"\0")
        else:
            pos = extra_off + len(extra)
            ifd += struct.pack("<HHII", tag, typ, count, pos)
            extra += raw + (b"\0" if len(raw) & 1 else b"")
    ifd += struct.pack("<I", 0)
    with open(out_file, "wb") as fp:
        fp.write(b"II" + struct.pack("<HI", 42, ifd_off))
        fp.write(data)
        if (8 + len(data)) & 1:
            fp.write(b"\0")
        fp.write(ifd)
        fp.write(extra)


def _read_tags(raw):
    if raw[:4] != b"II*\0":
        raise ValueError("not a little-endian classic TIFF")
    (off,) = struct.unpack_from("<I", raw, 4)
    (n,) = struct.unpack_from("<H", raw, off)
    tags = {}
    for k in range(n):
        tag, typ, count, val = struct.unpack_from("<HHII", raw, off + 2 + 12 * k)
        size = _SIZES.get(typ, 1) * count
        pos = off + 2 + 12 * k + 8 if size <= 4 else val
        blob = raw[pos:pos + size]
        if typ == _T_ASCII:
            tags[tag] = blob.rstrip(b"\0").decode(errors="replace")
        elif typ in (_T_SHORT, _T_LONG, _T_DOUBLE):
            fmt = {_T_SHORT: "H", _T_LONG: "I", _T_DOUBLE: "d"}[typ]
            tags[tag] = list(struct.unpack("<%d%s" % (count, fmt), blob))
    return tags


def read_dsm_tif(file):
    assert os.path.exists(file)
    raw = open(file, "rb").read()
    t = _read_tags(raw)
    w, h = t[256][0], t[257][0]
    if t.get(259, [1])[0] != 1 or t.get(339, [1])[0] != 3 or t.get(258, [0])[0] != 32:
        raise ValueError("only uncompressed float32 DSMs are read")
    image = np.concatenate([np.frombuffer(raw, dtype="<f4", count=c // 4, offset=o)
                            for o, c in zip(t[273], t[279])]).reshape(h, w).astype(np.float32)
    nodata = None
    if 42113 in t and t[42113].strip().lower() != "nan":
        nodata = float(t[42113])
        image[np.isclose(image, nodata)] = np.nan
    scale, tie = t[33550], t[33922]
    keys = t.get(34735, [])
    epsg = 0
    for k in range(keys[3] if keys else 0):
        if keys[4 + 4 * k] == 3072:
            epsg = keys[4 + 4 * k + 3]
    if 32601 <= epsg <= 32660:
        zone_number, hemisphere = epsg - 32600, "N"
    elif 32701 <= epsg <= 32760:
        zone_number, hemisphere = epsg - 32700, "S"
    else:
        zone_number, hemisphere = 0, "N"
    ul_e, ul_n = tie[3] - tie[0] * scale[0], tie[4] + tie[1] * scale[1]
    geo = (ul_e, scale[0], 0.0, ul_n, 0.0, -scale[1])
    meta = {"geo": geo, "proj": t.get(34737, "").rstrip("|"), "meta": {"AREA_OR_POINT": "Area"},
            "img_width": w, "img_height": h, "nodata": nodata,
            "zone_number": zone_number, "hemisphere": hemisphere,
            "ul_easting": geo[0], "ul_northing": geo[3],
            "east_resolution": geo[1], "north_resolution": abs(geo[5])}
    meta["lr_easting"] = meta["ul_easting"] + (w - 1) * meta["east_resolution"]
    meta["lr_northing"] = meta["ul_northing"] - (h - 1) * meta["north_resolution"]
    meta["area_width"] = meta["lr_easting"] - meta["ul_easting"]
    meta["area_height"] = meta["ul_northing"] - meta["lr_northing"]
    finite = image[~np.isnan(image)]
    meta["alt_min"] = float(finite.min()) if finite.size else float("nan")
    meta["alt_max"] = float(finite.max()) if finite.size else float("nan")
    return image, meta
