"""Film output: SORT's EXR file and the display tile protocols (host paths; no GPU).

The oracle's half conversion is pinned against numpy's IEEE conversion wherever the two rounding
rules agree and against closed forms where tinyexr's differs (exact ties round away from zero,
float denormals flush to zero); its EXR layout against an independent reader and a
hand-assembled header; the library's bytes must then equal the oracle's for films holding every
kind of value, and display packets must reassemble into the film the way the receivers place
them, over a real socket too.
"""
import socket
import struct
import threading

import numpy as np
import pytest

from oracle import film_oracle as O
from sort_amd import SortGpuError, film


def _film(w, h, seed):
    g = np.random.default_rng(seed)
    f = (g.random((h, w, 3)) * np.float32(4.0)).astype(np.float32)
    f[0, 0] = [np.inf, -np.inf, np.nan]
    f[0, 1] = [0.0, -0.0, 1e-40]          # zero, negative zero, a float denormal
    f[0, 2] = [65504.0, 65520.0, 1e9]      # largest half, the first value that overflows, far beyond
    f[0, 3] = [6.1e-5, 3e-7, 5.9604645e-8]  # half normal / denormal range
    f[0, 4] = [1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, -(2.0 + 2.0 ** -10)]  # exact ties
    f[1] = -f[1]
    return f


def test_half_conversion_pinned():
    g = np.random.default_rng(0)
    x = np.concatenate([g.normal(size=200000) * 10.0 ** g.integers(-9, 6, 200000),
                        g.uniform(-70000, 70000, 10000)]).astype(np.float32)
    got = O.half_tinyexr(x)
    with np.errstate(over="ignore"):
        ieee = x.astype(np.float16).view(np.uint16)
    # the dropped bits are exactly half an ulp of the half result: a tie
    b = x.view(np.uint32).astype(np.int64)
    exp = (b >> 23) & 0xFF
    newexp = exp - 127 + 15
    shift = np.where(newexp > 0, 13, np.clip(14 - newexp, 0, 40))
    mant = (b & 0x7FFFFF) | np.where(newexp > 0, 0, 0x800000)
    dropped = mant & ((np.int64(1) << shift) - 1)
    tie = dropped == (np.int64(1) << (shift - 1))
    float_denormal = exp == 0
    agree = ~tie & ~float_denormal & (newexp > -11)
    assert np.array_equal(got[agree], ieee[agree])
    # ties: magnitude rounds up (IEEE rounds to even)
    t = np.float32([1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, -(1.0 + 2.0 ** -11), 2.0 ** -15 + 2.0 ** -25])
    want = np.float16([1.0 + 2.0 ** -10, 1.0 + 2.0 ** -9, -(1.0 + 2.0 ** -10), 0.0]).view(np.uint16)
    want[3] = np.uint16(0x0200 + 1)  # half denormal 2^-15 = 0x200, plus half its 2^-24 ulp: rounds up
    assert np.array_equal(O.half_tinyexr(t), want)
    assert np.float32(2.0 ** -15 + 2.0 ** -25).astype(np.float16).view(np.uint16) == 0x0200  # IEEE: to even
    # float denormals flush to signed zero; overflow to infinity; NaN to a quiet NaN
    assert list(O.half_tinyexr(np.float32([1e-40, -1e-40, 65520.0, -1e9, np.nan, np.inf]))) == \
        [0x0000, 0x8000, 0x7C00, 0xFC00, 0x7E00, 0x7C00]


def test_reader_on_hand_assembled_file():
    # a 16 x 16 file with one FLOAT channel "Y", written field by field from the OpenEXR layout
    w = h = 16
    head = b"\x76\x2f\x31\x01\x02\x00\x00\x00"
    ch = b"Y\0" + struct.pack("<i", 2) + b"\0\0\0\0" + struct.pack("<ii", 1, 1) + b"\0"
    for name, typ, data in [("channels", "chlist", ch), ("compression", "compression", b"\0"),
                            ("dataWindow", "box2i", struct.pack("<4i", 0, 0, 15, 15))]:
        head += name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data
    head += b"\0"
    vals = np.arange(w * h, dtype=np.float32).reshape(h, w)
    body = b"".join(struct.pack("<iI", y, 4 * w) + vals[y].tobytes() for y in range(h))
    offs = b"".join(struct.pack("<Q", len(head) + 8 * h + y * (8 + 4 * w)) for y in range(h))
    r = O.read_exr(head + offs + body)
    assert (r["width"], r["height"]) == (16, 16) and np.array_equal(r["channels"]["Y"], vals)


@pytest.mark.parametrize("w,h", [(16, 16), (37, 21), (64, 48)])
def test_exr_bytes_equal_oracle_and_read_back(w, h, tmp_path):
    f = _film(w, h, w * h)
    data = film.encode_exr(f)
    ref = O.exr_file(f, w, h)
    assert data == ref
    r = O.read_exr(data)
    assert (r["width"], r["height"]) == (w, h)
    assert r["attributes"]["compression"][1] == b"\0"
    assert list(r["channels"]) == ["B", "G", "R"]
    for c, k in (("R", 0), ("G", 1), ("B", 2)):
        assert np.array_equal(r["half_bits"][c], O.half_tinyexr(f[:, :, k]))
    plain = np.isfinite(f) & (np.abs(f) < 60000) & (np.abs(f) > 1e-4)
    back = np.stack([r["channels"]["R"], r["channels"]["G"], r["channels"]["B"]], -1)
    assert np.all(np.abs(back[plain] - f[plain]) <= np.abs(f[plain]) * 2.0 ** -11)
    path = tmp_path / "out.exr"
    film.write_exr(path, f)
    assert path.read_bytes() == data
    assert len(data) == film.lib().sortgpu_exr_size(w, h)


def test_exr_refusals(tmp_path):
    with pytest.raises(SortGpuError, match="16x16"):
        film.encode_exr(np.zeros((15, 16, 3), np.float32))
    with pytest.raises(SortGpuError, match="cannot open"):
        film.write_exr(tmp_path / "missing" / "x.exr", np.zeros((16, 16, 3), np.float32))


def test_blender_and_tev_packets_reassemble():
    w, h = 150, 90
    f = _film(w, h, 7)
    f[~np.isfinite(f)] = 0.5
    tiles = [(x, y, min(64, w - x), min(64, h - y)) for y in range(0, h, 64) for x in range(0, w, 64)]
    stream = b"".join(film.display_packet(film.BLENDER, f, *t) for t in tiles)
    img = O.read_blender_packets(stream, w, h)
    assert np.array_equal(img[:, :, :3], f) and np.all(img[:, :, 3] == 1.0)
    stream = film.display_packet(film.TEV_CREATE, f, title="render") + b"".join(
        film.display_packet(film.TEV, f, *t, title="render") for t in tiles)
    r = O.read_tev_packets(stream)
    assert r["create"] == ("render", w, h, ["R", "G", "B"], 1)
    for c, k in (("R", 0), ("G", 1), ("B", 2)):
        assert np.array_equal(r["images"]["render"][c], f[:, :, k])
    with pytest.raises(SortGpuError, match="outside"):
        film.display_packet(film.BLENDER, f, 140, 0, 64, 64)


def _server(mode, nbytes_box):
    srv = socket.socket(socket.AF_INET, socket.SOCK_STREAM)
    srv.bind(("127.0.0.1", 0))
    srv.listen(1)
    port = srv.getsockname()[1]

    def run():
        conn, _ = srv.accept()
        chunks = []
        while True:
            b = conn.recv(1 << 16)
            if not b:
                break
            chunks.append(b)
        nbytes_box.append(b"".join(chunks))
        conn.close()
        srv.close()

    t = threading.Thread(target=run, daemon=True)
    t.start()
    return port, t


@pytest.mark.parametrize("mode", [film.BLENDER, film.TEV])
def test_display_socket(mode):
    w, h = 100, 70
    f = _film(w, h, 9)
    f[~np.isfinite(f)] = 0.25
    box = []
    port, t = _server(mode, box)
    d = film.Display("127.0.0.1", port)
    d.send_film(f, mode=mode, title="live", tile=32)
    d.close()
    t.join(10)
    got = box[0]
    if mode == film.BLENDER:
        assert np.array_equal(O.read_blender_packets(got, w, h)[:, :, :3], f)
    else:
        r = O.read_tev_packets(got)
        assert r["create"][:3] == ("live", w, h)
        for c, k in (("R", 0), ("G", 1), ("B", 2)):
            assert np.array_equal(r["images"]["live"][c], f[:, :, k])
    with pytest.raises(SortGpuError, match="cannot connect"):
        film.Display("127.0.0.1", port)  # nothing listens any more
