"""Host hit reconstruction (sortgpu_hit_surface: setupIntersection, simd/simd_triangle.h:311-341,
and Quad::GetIntersect's hit fields, shape/quad.cpp:84-91) pinned by closed forms: a triangle in
a coordinate plane with axis-aligned attributes, the barycentric corners, the point/plane
relation, and transformed quads whose normal and tangent follow from the matrix."""
import ctypes
import math

import numpy as np
import pytest

from sort_amd import Obvh, SortGpuError, lib
from sort_amd._lib import Surface, check


def _surface(acc, ray, hit, normals=None, tangents=None, uv=None):
    r = np.ascontiguousarray(ray, np.float32)
    h = np.ascontiguousarray(hit, np.float32)
    keep = [np.ascontiguousarray(a, np.float32) if a is not None else None for a in (normals, tangents, uv)]
    ptr = [a.ctypes.data if a is not None else None for a in keep]
    out = Surface()
    check(lib().sortgpu_hit_surface(acc.handle, r.ctypes.data, h.ctypes.data, ptr[0], ptr[1], ptr[2],
                                    ctypes.byref(out)), acc.ctx.handle)
    return out


def _hit(t, prim, u, v):
    h = np.array([t, 0.0, u, v], np.float32)
    h[1:2].view(np.uint32)[0] = prim
    return h


def test_triangle_closed_forms(host_ctx):
    # triangle in z = 2: p0 = (0,0,2), p1 = (1,0,2), p2 = (0,1,2); counter-clockwise seen from +z
    pos = np.array([[0, 0, 2], [1, 0, 2], [0, 1, 2]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    acc = Obvh(host_ctx).build(pos, idx)
    normals = np.array([[0, 0, 1]] * 3, np.float32)
    tangents = np.array([[3, 0, 0], [3, 0, 0], [3, 0, 0]], np.float32)  # normalised by the helper
    uv = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]], np.float32)
    ray = np.array([0.25, 0.5, 0.0, 0.0, 0.0, 0.0, 1.0, 1e30], np.float32)
    s = _surface(acc, ray, _hit(2.0, 0, 0.25, 0.5), normals, tangents, uv)
    assert list(s.intersect) == [0.25, 0.5, 2.0] and s.t == 2.0
    # cross(p2 - p0, p1 - p0) = (0,1,0) x (1,0,0) = (0,0,-1): SORT's geometric normal faces -z here
    assert list(s.gnormal) == [0.0, 0.0, -1.0]
    assert list(s.normal) == [0.0, 0.0, 1.0] and list(s.tangent) == [1.0, 0.0, 0.0]
    assert list(s.view) == [-0.0, -0.0, -1.0]
    assert (s.u, s.v) == (0.25, 0.5)  # the texture coordinates of this triangle are its barycentrics
    assert s.prim == 0 and s.is_quad == 0


@pytest.mark.parametrize("corner,uvw", [(0, (0.0, 0.0)), (1, (1.0, 0.0)), (2, (0.0, 1.0))])
def test_triangle_corners_take_the_vertex_attributes(host_ctx, corner, uvw):
    g = np.random.default_rng(corner)
    pos = g.random((3, 3)).astype(np.float32)
    acc = Obvh(host_ctx).build(pos, np.array([[0, 1, 2]], np.uint32))
    nrm = g.normal(size=(3, 3)).astype(np.float32)
    tan = g.normal(size=(3, 3)).astype(np.float32)
    uv = g.random((3, 2)).astype(np.float32)
    ray = np.array([0, 0, 0, 0, 1, 1, 1, 1e30], np.float32)
    s = _surface(acc, ray, _hit(1.0, 0, *uvw), nrm, tan, uv)
    # at a corner the interpolation weights are exactly 1, 0, 0: the vertex's own (normalised) values
    n = nrm[corner].astype(np.float64)
    assert np.allclose(s.normal, n / np.linalg.norm(n), atol=1e-6)
    t = tan[corner].astype(np.float64)
    assert np.allclose(s.tangent, t / np.linalg.norm(t), atol=1e-6)
    assert np.array_equal(np.array([s.u, s.v], np.float32), uv[corner])


def test_intersect_lies_on_the_triangle(host_ctx):
    g = np.random.default_rng(5)
    pos = g.random((3, 3)).astype(np.float32)
    acc = Obvh(host_ctx).build(pos, np.array([[0, 1, 2]], np.uint32))
    u, v = np.float32(0.2), np.float32(0.3)
    target = (1 - u - v) * pos[0].astype(np.float64) + u * pos[1] + v * pos[2]
    o = np.array([-1.0, -2.0, 3.0])
    d = target - o
    ray = np.array([*o, 0.0, *d, 1e30], np.float32)
    s = _surface(acc, ray, _hit(1.0, 0, u, v))
    assert np.allclose(s.intersect, target, atol=1e-5)
    gn = np.cross(pos[2] - pos[0], pos[1] - pos[0]).astype(np.float64)
    assert np.allclose(s.gnormal, gn / np.linalg.norm(gn), atol=1e-6)
    assert list(s.normal) == [0.0, 0.0, 0.0] and (s.u, s.v) == (0.0, 0.0)  # attributes not given


def test_quad_normal_and_tangent_follow_the_transform(host_ctx):
    # light-to-world: rotate 90 degrees about x (local +y -> world +z, local +z -> world -y), then
    # translate; the shading normal is the transformed up vector, the tangent the transformed +z
    c, s_ = 0.0, 1.0
    mat = np.array([[1, 0, 0, 0.5], [0, c, -s_, 0.25], [0, s_, c, 2.0], [0, 0, 0, 1]], np.float32)
    inv = np.linalg.inv(mat.astype(np.float64)).astype(np.float32)
    quad = {"mat": mat, "inv": inv, "size_x": 1.0, "size_y": 1.0}
    pos = np.array([[5, 5, 5], [6, 5, 5], [5, 6, 5]], np.float32)
    acc = Obvh(host_ctx).build(pos, np.array([[0, 1, 2]], np.uint32), [quad])
    ray = np.array([0.5, 0.25, 0.0, 0.0, 0.0, 0.0, 1.0, 1e30], np.float32)
    s = _surface(acc, ray, _hit(2.0, 1, 0.0, 0.0))
    assert s.is_quad == 1 and s.prim == 1
    assert np.allclose(s.normal, [0, 0, 1], atol=1e-6) and list(s.gnormal) == list(s.normal)
    assert np.allclose(s.tangent, [0, -1, 0], atol=1e-6)
    assert list(s.intersect) == [0.5, 0.25, 2.0]


def test_refuses_misses_and_foreign_ids(host_ctx):
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    acc = Obvh(host_ctx).build(pos, np.array([[0, 1, 2]], np.uint32))
    ray = np.zeros(8, np.float32)
    for prim in (0xFFFFFFFF, 1, 7):
        with pytest.raises(SortGpuError) as e:
            _surface(acc, ray, _hit(1.0, prim, 0.0, 0.0))
        assert e.value.code == -1
