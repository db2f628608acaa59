"""The oracle's random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
numbers: as easy as 1, 2, 3", SC'11) against the known-answer vectors the authors publish with
Random123 (kat_vectors, philox4x32 with 10 rounds), and the 24-bit float mapping of SORT's
sort_rand<float> (core/rand.cpp: the top 24 bits of a draw over 2^24, so every value lies in
[0, 1) on a grid of 2^-24). The product draws with the same function and keys (renders with the
same keys agree path by path), so these vectors pin both."""
import ctypes

import numpy as np
import pytest

from oracle.oracle import lib as oracle_lib

KAT = [
    # counter (4 words), key (2 words), expected output (4 words)
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


def _lib():
    l = oracle_lib()
    l.oracle_philox.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
    l.oracle_draws.argtypes = [ctypes.c_uint32] * 5 + [ctypes.c_void_p]
    return l


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox4x32_10_known_answers(ctr, key, want):
    c = np.array(ctr, np.uint32)
    out = np.zeros(4, np.uint32)
    _lib().oracle_philox(c.ctypes.data, key[0], key[1], out.ctypes.data)
    assert [int(x) for x in out] == list(want)


def test_draws_are_philox_of_the_key_mapped_to_24_bits():
    l = _lib()
    g = np.random.default_rng(3)
    for _ in range(200):
        seed, pixel, sample, bounce, call = (int(x) for x in g.integers(0, 2**32, 5, dtype=np.uint64))
        c = np.array([sample, bounce, call, 0], np.uint32)
        raw = np.zeros(4, np.uint32)
        l.oracle_philox(c.ctypes.data, seed, pixel, raw.ctypes.data)
        d = np.zeros(4, np.float32)
        l.oracle_draws(seed, pixel, sample, bounce, call, d.ctypes.data)
        want = (raw >> 8).astype(np.float64) / 2.0**24
        assert np.array_equal(d.astype(np.float64), want)
        assert np.all(d >= 0.0) and np.all(d < 1.0)


def test_mapping_extremes():
    # the largest draw maps to 1 - 2^-24 (never 1), the smallest to 0
    assert np.float32((0xFFFFFFFF >> 8) * np.float32(1.0 / 16777216.0)) == np.float32(1.0 - 2.0**-24)
    assert (0x000000FF >> 8) == 0


def test_camera_rays_closed_form():
    """PerspectiveCamera::GenerateRay (camera/perspective.cpp:26-104) for a pinhole camera whose up
    vector is perpendicular to the view direction, as the exporter writes it (its up is a row of
    the orthonormal camera matrix, exporter.py:167-169): the ray of raster point (x + ju, y + jv)
    leaves the eye along cx / xScale * right + cy / yScale * up + forward with
    cx = 2 (x + ju) / W - 1, cy = 1 - 2 (y + jv) / H, yScale = 1 / tan(fov / 2), xScale =
    yScale / aspect. Computed here in float64 from the jitter draws and compared with the
    oracle's recorded camera rays."""
    import math

    from oracle.oracle import OraclePT
    from sort_amd import scenes
    sc = scenes.cornell_sphere()
    fwd = np.array([0.3, -0.2, 1.0])
    fwd /= np.linalg.norm(fwd)
    up = np.cross(fwd, np.cross(np.array([0.0, 1.0, 0.0]), fwd))
    up /= np.linalg.norm(up)
    eye = np.array([0.45, 0.55, 0.05])
    cam = dict(eye=tuple(eye), target=tuple(eye + fwd), up=tuple(up), fov_y_deg=55.0)
    W, H, spp, seed = 24, 16, 2, 3
    rays, _, valid = OraclePT(sc).ray_dump(cam, W, H, spp, 1, seed, 0, W * H, threads=2)
    assert valid.all()
    l = _lib()
    right = np.cross(up.astype(np.float32), fwd.astype(np.float32)).astype(np.float64)
    right /= np.linalg.norm(right)
    up2 = np.cross(fwd, right)
    y_scale = 1.0 / math.tan(math.radians(55.0) * 0.5)
    x_scale = y_scale / (W / H)
    d = np.zeros(4, np.float32)
    for pix in range(W * H):
        x, y = pix % W, pix // W
        for s in range(spp):
            l.oracle_draws(seed, pix, s, 0xFFFF, 0, d.ctypes.data)
            cx = 2.0 * (x + float(d[0])) / W - 1.0
            cy = 1.0 - 2.0 * (y + float(d[1])) / H
            want = cx / x_scale * right + cy / y_scale * up2 + fwd
            want /= np.linalg.norm(want)
            r = rays[pix * spp + s]
            assert np.allclose(r[0:3], eye, atol=1e-6) and r[3] == 0.0
            got = r[4:7].astype(np.float64)
            assert np.allclose(got / np.linalg.norm(got), want, atol=2e-6), (pix, s)
