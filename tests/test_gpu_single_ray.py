"""Single-ray compatibility forms (sortgpu_trace_closest1 / _any1, SORT's per-ray GetIntersect and
IsOccluded): every result bitwise equal to the batched calls' (which are bitwise the oracle's),
called from several threads at once as SORT's workers call them, and hit reconstruction from a
single-ray hit matching the path tracer's own surface normal at that point."""
import ctypes
import threading

import numpy as np
import pytest
import torch

from oracle.oracle import OracleBvh
from sort_amd import Obvh, lib, scenes
from sort_amd._lib import Surface, check

pytestmark = pytest.mark.gpu


def _closest1(acc, ray):
    r = np.ascontiguousarray(ray, np.float32)
    h = np.zeros(4, np.float32)
    check(lib().sortgpu_trace_closest1(acc.handle, r.ctypes.data, h.ctypes.data), acc.ctx.handle)
    return h


def _any1(acc, ray):
    r = np.ascontiguousarray(ray, np.float32)
    occ = ctypes.c_int(-1)
    check(lib().sortgpu_trace_any1(acc.handle, r.ctypes.data, ctypes.byref(occ)), acc.ctx.handle)
    return occ.value


def test_single_ray_equals_batched_from_many_threads(gpu_ctx):
    sc = scenes.cornell_sphere()
    acc = Obvh(gpu_ctx).build(sc["pos"], sc["idx"], sc["quads"])
    rays = np.concatenate([scenes.random_rays(600, 31, lo=0.05, hi=0.95),
                           scenes.pinhole_rays(16, 16, eye=(0.5, 0.5, 0.02), target=(0.5, 0.5, 1.0), vfov_deg=70.0)])
    ref = acc.get_intersect(rays)
    ref_occ = acc.is_occluded(rays)
    ob = OracleBvh(sc["pos"], sc["idx"], sc["quads"])
    assert np.array_equal(ref.view(np.uint32), ob.closest(rays).view(np.uint32))
    got = np.zeros_like(ref)
    got_occ = np.full(len(rays), -1, np.int32)
    errors = []

    def worker(k):
        try:
            for i in range(k, len(rays), 6):
                got[i] = _closest1(acc, rays[i])
                got_occ[i] = _any1(acc, rays[i])
        except Exception as e:  # reported below
            errors.append(e)

    print(f"[single] {len(rays)} rays from 6 threads", flush=True)
    th = [threading.Thread(target=worker, args=(k,)) for k in range(6)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(got_occ, ref_occ.astype(np.int32))
    assert (ref[:, 1].view(np.uint32) != 0xFFFFFFFF).mean() > 0.9


def test_single_ray_hit_reconstruction(gpu_ctx):
    sc = scenes.cornell_sphere()
    acc = Obvh(gpu_ctx).build(sc["pos"], sc["idx"], sc["quads"])
    normals = np.ascontiguousarray(sc["normals"], np.float32)
    ray = np.array([0.5, 0.5, 0.02, 0.0, 0.0, -0.05, 1.0, 3.4e38], np.float32)  # towards the sphere
    h = _closest1(acc, ray)
    prim = int(h[1:2].view(np.uint32)[0])
    assert prim < len(sc["idx"])
    s = Surface()
    check(lib().sortgpu_hit_surface(acc.handle, ray.ctypes.data, h.ctypes.data, normals.ctypes.data, None, None,
                                    ctypes.byref(s)), acc.ctx.handle)
    p = np.array(s.intersect)
    assert np.allclose(p, ray[:3] + h[0] * ray[4:7], atol=1e-6)
    n = np.array(s.normal)
    assert abs(np.linalg.norm(n) - 1.0) < 1e-5 and np.dot(n, ray[4:7]) < 0  # faces the ray
