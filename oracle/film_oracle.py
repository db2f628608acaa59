"""Test infrastructure: film output formats, restated and read back independently.

* ``half_tinyexr``: tinyexr's float_to_half_full (thirdparty/tiny_exr/tinyexr.h:7133-7171, the
  ISPC conversion), vectorised over float32 bit patterns.
* ``exr_file``: the bytes of tinyexr's SaveEXR(rgb, w, h, 3, save_as_fp16 = 1) as SORT calls it
  (texture/texturebase.cpp:27-57; SaveEXR tinyexr.h:12854-12990 and SaveEXRImageToMemory
  :11439-11830: magic, version 2, channels B G R half, compression 0, data and display windows,
  lineOrder 0, pixelAspectRatio 1, screenWindowCenter (0, 0), screenWindowWidth w, offset table,
  one uncompressed scanline per chunk).
* ``read_exr``: a generic reader of single-part scanline EXR files with uncompressed HALF/FLOAT
  channels, written from the OpenEXR file layout, not from the writer above.
* ``read_blender_packets`` / ``read_tev_packets``: the receiving side of the display protocols
  (blender-plugin/addons/sortblend/renderer.py:31-86 for Blender; tev's UpdateImage and
  CreateImage packets as core/display_mgr.cpp:86-140 sends them).

Only tests/ may use this module. Pinned by tests/test_film_out.py: the half conversion against
numpy's IEEE round-to-nearest-even conversion wherever the two rules agree (every value that is
not exactly halfway between two halves) and against closed forms where they differ (ties,
float denormals, overflow, NaN); the EXR bytes against the reader; the reader against a
hand-assembled file.
"""
from __future__ import annotations

import struct

import numpy as np


def half_tinyexr(x) -> np.ndarray:
    b = np.asarray(x, np.float32).view(np.uint32).astype(np.int64)
    sign = b >> 31
    exp = (b >> 23) & 0xFF
    mant = b & 0x7FFFFF
    o = np.zeros_like(b)
    special = exp == 255
    o = np.where(special, (31 << 10) | np.where(mant != 0, 0x200, 0), o)
    newexp = exp - 127 + 15
    normal = (exp != 0) & ~special
    over = normal & (newexp >= 31)
    o = np.where(over, 31 << 10, o)
    under = normal & (newexp <= 0) & (14 - newexp <= 24)
    m = mant | 0x800000
    sh = np.clip(14 - newexp, 0, 63)
    sh1 = np.clip(13 - newexp, 0, 63)
    ou = ((m >> sh) & 0x3FF) + ((m >> sh1) & 1)
    o = np.where(under, ou, o)
    mid = normal & (newexp > 0) & (newexp < 31)
    om = ((np.clip(newexp, 0, 31) << 10) | (mant >> 13)) + ((mant >> 12) & 1)
    o = np.where(mid, om, o)
    return ((o & 0x7FFF) | (sign << 15)).astype(np.uint16)


def _attr(name, typ, data):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data


def exr_file(rgb, w, h) -> bytes:
    rgb = np.asarray(rgb, np.float32).reshape(h, w, 3)
    ch = b""
    for name in ("B", "G", "R"):
        ch += name.encode() + b"\0" + struct.pack("<i", 1) + b"\0\0\0\0" + struct.pack("<ii", 1, 1)
    ch += b"\0"
    head = b"\x76\x2f\x31\x01" + b"\x02\x00\x00\x00"
    head += _attr("channels", "chlist", ch)
    head += _attr("compression", "compression", b"\0")
    box = struct.pack("<4i", 0, 0, w - 1, h - 1)
    head += _attr("dataWindow", "box2i", box) + _attr("displayWindow", "box2i", box)
    head += _attr("lineOrder", "lineOrder", b"\0")
    head += _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    head += _attr("screenWindowCenter", "v2f", struct.pack("<2f", 0.0, 0.0))
    head += _attr("screenWindowWidth", "float", struct.pack("<f", float(w)))
    head += b"\0"
    chunk = 8 + 6 * w
    offsets = b"".join(struct.pack("<Q", len(head) + 8 * h + y * chunk) for y in range(h))
    halves = half_tinyexr(rgb)  # (h, w, 3)
    body = bytearray()
    for y in range(h):
        body += struct.pack("<iI", y, 6 * w)
        for c in (2, 1, 0):  # B, G, R
            body += halves[y, :, c].astype("<u2").tobytes()
    return head + offsets + bytes(body)


def read_exr(data: bytes) -> dict:
    """{'attributes': {name: (type, bytes)}, 'width', 'height', 'channels': {name: float32 (h, w)},
    'half_bits': {name: uint16 (h, w)} for HALF channels}."""
    if data[:4] != b"\x76\x2f\x31\x01":
        raise ValueError("not an OpenEXR file")
    version = struct.unpack_from("<I", data, 4)[0]
    if version & 0xFF != 2 or version & 0x200:
        raise ValueError("only single-part scanline files")
    pos = 8
    attrs = {}
    while data[pos] != 0:
        e = data.index(b"\0", pos)
        name = data[pos:e].decode()
        e2 = data.index(b"\0", e + 1)
        typ = data[e + 1:e2].decode()
        size = struct.unpack_from("<i", data, e2 + 1)[0]
        attrs[name] = (typ, data[e2 + 5:e2 + 5 + size])
        pos = e2 + 5 + size
    pos += 1
    chl = attrs["channels"][1]
    chans, p = [], 0
    while chl[p] != 0:
        e = chl.index(b"\0", p)
        name = chl[p:e].decode()
        ptype, _, xs, ys = struct.unpack_from("<iBxxxii", chl, e + 1)
        chans.append((name, ptype))
        p = e + 17
    if attrs["compression"][1] != b"\0":
        raise ValueError("compressed files are not read here")
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f"<{h}Q", data, pos)
    size = {1: 2, 2: 4}
    out = {n: np.zeros((h, w), np.float32) for n, _ in chans}
    bits = {n: np.zeros((h, w), np.uint16) for n, t in chans if t == 1}
    for off in offsets:
        y, n = struct.unpack_from("<iI", data, off)
        q = off + 8
        for name, t in chans:  # channels in the file's (alphabetical) order
            raw = data[q:q + size[t] * w]
            if t == 1:
                hb = np.frombuffer(raw, "<u2")
                bits[name][y - y0] = hb
                out[name][y - y0] = hb.view(np.float16).astype(np.float32)
            else:
                out[name][y - y0] = np.frombuffer(raw, "<f4")
            q += size[t] * w
        if q - off - 8 != n:
            raise ValueError("chunk size mismatch")
    return {"attributes": attrs, "width": w, "height": h, "channels": out, "half_bits": bits}


def read_blender_packets(stream: bytes, width: int, height: int) -> np.ndarray:
    """Reassembles the film (top row first, RGBA) from Blender-mode tile packets the way the
    add-on's receiver places them (render_engine.begin_result(x, H - y - h, w, h) with rows from
    the tile's bottom up)."""
    img = np.full((height, width, 4), np.nan, np.float32)
    pos = 0
    while pos < len(stream):
        length = int.from_bytes(stream[pos:pos + 4], "little")
        tw, th, x, y = struct.unpack_from("<4i", stream, pos + 4)
        pix = np.frombuffer(stream[pos + 20:pos + 4 + length], np.float32).reshape(th, tw, 4)
        by = height - y - th  # Blender's bottom-up row of the tile
        bl = np.flipud(img)   # Blender's bottom-up view of the image
        bl[by:by + th, x:x + tw] = pix
        pos += 4 + length
    return img


def read_tev_packets(stream: bytes) -> dict:
    """{'create': (title, w, h, channels) or None, 'images': {title: {channel: float32 (h, w)}}}."""
    res = {"create": None, "images": {}}
    pos = 0
    while pos < len(stream):
        length = struct.unpack_from("<i", stream, pos)[0]
        kind, grab = stream[pos + 4], stream[pos + 5]
        q = pos + 6
        e = stream.index(b"\0", q)
        title = stream[q:e].decode()
        q = e + 1
        if kind == 4:
            w, h, n = struct.unpack_from("<3i", stream, q)
            q += 12
            names = []
            for _ in range(n):
                e = stream.index(b"\0", q)
                names.append(stream[q:e].decode())
                q = e + 1
            res["create"] = (title, w, h, names, grab)
            res["images"][title] = {c: np.full((h, w), np.nan, np.float32) for c in names}
        elif kind == 3:
            e = stream.index(b"\0", q)
            chan = stream[q:e].decode()
            q = e + 1
            x, y, w, h = struct.unpack_from("<4i", stream, q)
            q += 16
            px = np.frombuffer(stream[q:q + 4 * w * h], np.float32).reshape(h, w)
            res["images"][title][chan][y:y + h, x:x + w] = px
            q += 4 * w * h
        else:
            raise ValueError(f"unknown packet type {kind}")
        if q != pos + length:
            raise ValueError("packet length mismatch")
        pos += length
    return res
