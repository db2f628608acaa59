"""Pins the oracle's arithmetic primitives with closed forms and the reference's own SIMD
known answers (unittests/simd.hpp). Each case is chosen so a plausible restatement mistake
(wrong axis permutation, FMA contraction, det order, open/closed window, wrong min/max
operand) changes the answer."""
import math

import numpy as np
import pytest

from oracle import oracle as O

FLT_MAX = np.float32(3.4028234663852886e38)


def ray(o, d, tmin=0.0, tmax=FLT_MAX):
    return np.array([*o, tmin, *d, tmax], np.float32)


# --- SIMD wrapper semantics (unittests/simd.hpp:180-197, 484-512) ---------------------------

def test_rcp_is_ieee_division_bitwise():
    x = np.array([3.0, 7.0, 0.1, -1e-3, 1e30, 2.5e-38, 123.456, -9.0], np.float32)
    got = O.simd_rcp8(x)
    want = (np.float32(1.0) / x).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_minreduction_and_movemask():
    x = np.array([5, 3, 9, 2.5, 7, 2.25, 8, 4], np.float32)
    assert O.simd_minreduce8(x) == np.float32(2.25)
    x = np.array([-1, 2, -3, 4, -5, 6, -7, 8], np.float32)
    assert O.simd_movemask8(x) == 0b01010101


# --- Ray::Prepare / resolveRayData (math/ray.h:60-69, simd_ray_utils.h:60-83) ---------------

@pytest.mark.parametrize("d,k", [((0, 0, 1), (1, 2, 0)), ((0, 1, 0), (0, 1, 2)),
                                 ((1, 0, 0), (2, 0, 1)),
                                 ((0.6, 0.6, 0.5), (0, 1, 2)),   # |x| == |y|: x not strictly largest -> y
                                 ((0.5, 0.5, 0.5), (1, 2, 0)),   # all equal -> z
                                 ((-3, 1, 2), (2, 0, 1)),        # sign ignored
                                 ])
def test_prepare_major_axis(d, k):
    kk, _ = O.ray_setup(ray((0, 0, 0), d))
    assert tuple(int(x) for x in kk) == k


def test_prepare_scales_closed_form():
    d = np.array([0.25, -0.5, 2.0], np.float32)  # major z -> ky=2, kz=0, kx=1
    k, f = O.ray_setup(ray((1, 2, 3), d))
    assert list(k) == [1, 2, 0]
    sx, sy, sz = f[0:3]
    assert sx == np.float32(-d[1]) / d[2] and sz == np.float32(-d[0]) / d[2] and sy == np.float32(1) / d[2]
    # child-box reciprocals and -o/d (no clamp needed here)
    assert np.array_equal(f[3:6], (np.float32(1) / d).astype(np.float32))
    assert np.array_equal(f[6:9], (-np.array([1, 2, 3], np.float32) / d).astype(np.float32))


def test_small_direction_clamp():
    # |d_i| < 1e-5 -> sign(d_i) * 1e-5 with sign(0) = +1, only for the child-box constants
    d = np.array([0.0, -5e-6, 1.0], np.float32)
    _, f = O.ray_setup(ray((0.5, 0.5, 0.5), d))
    assert f[3] == np.float32(1) / np.float32(1e-5)
    assert f[4] == np.float32(1) / np.float32(-1e-5)
    assert f[6] == np.float32(-0.5) / np.float32(1e-5)


# --- watertight triangle test (simd_triangle.h:212-300) -------------------------------------

def test_triangle_axis_aligned_exact_t():
    tri = np.array([0, 0, 5, 1, 0, 5, 0, 1, 5], np.float32)  # plane z = 5
    ok, tuv = O.tri_simd(ray((0.25, 0.25, 0), (0, 0, 1)), tri)
    assert ok and tuv[0] == np.float32(5.0)
    # barycentrics of (0.25, 0.25) in (v0, v1, v2): u = 0.25 (v1), v = 0.25 (v2)
    assert abs(tuv[1] - 0.25) < 1e-6 and abs(tuv[2] - 0.25) < 1e-6


def test_triangle_window_is_half_open():
    tri = np.array([0, 0, 5, 1, 0, 5, 0, 1, 5], np.float32)
    r = ray((0.25, 0.25, 0), (0, 0, 1), tmin=0.0, tmax=5.0)
    assert O.tri_simd(r, tri)[0]                     # t == tmax accepted (cmple)
    r = ray((0.25, 0.25, 0), (0, 0, 1), tmin=5.0, tmax=10.0)
    assert not O.tri_simd(r, tri)[0]                 # t == tmin rejected (cmpgt)
    # the scalar test is open at both ends (triangle.cpp:89)
    assert not O.tri_scalar(ray((0.25, 0.25, 0), (0, 0, 1), 0.0, 5.0), tri)[0]


def test_triangle_both_windings_and_miss():
    tri = np.array([0, 0, 5, 0, 1, 5, 1, 0, 5], np.float32)  # reversed winding still hits
    assert O.tri_simd(ray((0.25, 0.25, 0), (0, 0, 1)), tri)[0]
    assert not O.tri_simd(ray((0.75, 0.75, 0), (0, 0, 1)), tri)[0]
    assert not O.tri_simd(ray((0.25, 0.25, 6), (0, 0, 1)), tri)[0]  # behind the origin


def test_triangle_degenerate_rejected():
    tri = np.array([0, 0, 5, 1, 1, 5, 2, 2, 5], np.float32)  # collinear: det == 0
    assert not O.tri_simd(ray((1, 1, 0), (0, 0, 1)), tri)[0]


def test_watertight_shared_edge_never_misses_both():
    # two triangles sharing the diagonal of the unit square at z = 1; rays through the edge
    a = np.array([0, 0, 1, 1, 0, 1, 1, 1, 1], np.float32)
    b = np.array([0, 0, 1, 1, 1, 1, 0, 1, 1], np.float32)
    g = np.random.default_rng(3)
    for s in g.uniform(0.05, 0.95, 400).astype(np.float32):
        d = g.normal(size=3).astype(np.float32)
        d[2] = abs(d[2]) + 0.2
        d /= np.linalg.norm(d)
        e = np.array([s, s, 1.0], np.float32)          # a point of the shared edge
        o = (e - np.float32(2.0) * d).astype(np.float32)
        r = ray(o, d)
        assert O.tri_simd(r, a)[0] or O.tri_simd(r, b)[0]


def test_simd_vs_scalar_det_order_can_differ_only_in_last_ulps():
    g = np.random.default_rng(5)
    for _ in range(200):
        tri = g.random(9).astype(np.float32)
        o = g.random(3).astype(np.float32) - 0.5
        c = tri.reshape(3, 3).mean(0)
        d = (c - o) / np.linalg.norm(c - o)
        r = ray(o, d.astype(np.float32))
        a, b = O.tri_simd(r, tri), O.tri_scalar(r, tri)
        if a[0] and b[0]:
            assert abs(a[1][0] - b[1][0]) <= 4 * np.spacing(np.float32(abs(b[1][0])))


# --- slab tests (simd_bbox.h:62-85, bbox.h:124-156) -----------------------------------------

def test_bbox8_closed_form_entry_distance():
    mn = np.zeros((3, 8), np.float32)
    mx = np.ones((3, 8), np.float32)
    for j in range(8):
        mn[2, j] = 2.0 + j
        mx[2, j] = 3.0 + j
    valid = [1, 1, 1, 1, 1, 1, 1, 0]
    m, f = O.bbox8(ray((0.5, 0.5, 0), (0, 0, 1)), mn, mx, valid)
    assert m == 0x7F
    assert list(f[:7]) == [2.0 + j for j in range(7)] and f[7] == -1.0


def test_bbox8_respects_tmax_and_miss_sets_minus_one():
    mn = np.zeros((3, 8), np.float32)
    mx = np.ones((3, 8), np.float32)
    mn[2, :] = 10.0
    mx[2, :] = 11.0
    m, f = O.bbox8(ray((0.5, 0.5, 0), (0, 0, 1), 0.0, 9.0), mn, mx, [1] * 8)
    assert m == 0 and np.all(f == -1.0)
    m, _ = O.bbox8(ray((0.5, 0.5, 0), (0, 0, 1), 0.0, 10.0), mn, mx, [1] * 8)
    assert m == 0xFF  # fmin <= fmax (touching)


def test_bbox8_contains_origin_gives_tmin():
    mn = -np.ones((3, 8), np.float32)
    mx = np.ones((3, 8), np.float32)
    m, f = O.bbox8(ray((0, 0, 0), (0.3, 0.4, 0.5), 0.25, 100.0), mn, mx, [1] * 8)
    assert m == 0xFF and np.all(f == np.float32(0.25))


def test_root_slab_parallel_axis_branch():
    box = np.array([0, 0, 0, 1, 1, 1], np.float32)
    assert O.root_slab(ray((0.5, 0.5, -1), (0, 0, 1)), box) == np.float32(1.0)
    assert O.root_slab(ray((2, 0.5, -1), (0, 0, 1)), box) == -1.0   # outside on a parallel axis
    assert O.root_slab(ray((0.5, 0.5, -1), (0, 0, -1)), box) == -1.0


# --- quad (shape/quad.cpp:66-94) ------------------------------------------------------------

def _quad(center, size):
    from sort_amd.scenes import translate_quad
    return translate_quad(center, size, facing_down=True)


def test_quad_hit_and_bounds():
    q = _quad((0.5, 1.0, 0.5), 0.5)
    ok, t = O.quad(ray((0.5, 0, 0.5), (0, 1, 0)), q)
    assert ok and t == np.float32(1.0)
    assert not O.quad(ray((0.8, 0, 0.5), (0, 1, 0)), q)[0]          # |x| > half size
    assert O.quad(ray((0.75, 0, 0.5), (0, 1, 0)), q)[0]             # on the edge: inclusive
    assert not O.quad(ray((0.5, 0, 0.5), (1, 0, 0)), q)[0]          # parallel
    assert not O.quad(ray((0.5, 0, 0.5), (0, 1, 0), 1.0, 5.0), q)[0]  # t == tmin rejected
    assert O.quad(ray((0.5, 0, 0.5), (0, 1, 0), 0.0, 1.0), q)[0]      # t == tmax accepted
    # two-sided for intersection (emission is one-sided, not the test)
    assert O.quad(ray((0.5, 2, 0.5), (0, -1, 0)), q)[0]
