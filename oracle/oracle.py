"""ctypes binding of the oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/liboracle.so — the CPU restatement of SORT's OBVH/QBVH (obvh_oracle.cpp) — for the
parity tests and for bench.py's CPU-baseline leg. The product package never imports this.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None
_P = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = ctypes.CDLL(LIB_PATH)
        l.oracle_build.restype = _P
        l.oracle_build.argtypes = [_P, ctypes.c_uint32, _P, ctypes.c_uint32, _P, ctypes.c_uint32, ctypes.c_int,
                                   ctypes.c_uint32, ctypes.c_uint32]
        l.oracle_destroy.argtypes = [_P]
        l.oracle_info.argtypes = [_P, _P]
        l.oracle_scene_bbox.argtypes = [_P, _P]
        l.oracle_prim_order.argtypes = [_P, _P]
        l.oracle_tree_dump.restype = ctypes.c_uint64
        l.oracle_tree_dump.argtypes = [_P, _P, ctypes.c_uint64]
        l.oracle_trace.restype = ctypes.c_double
        l.oracle_trace.argtypes = [_P, _P, ctypes.c_uint64, _P, _P, ctypes.c_int, ctypes.c_int, _P]
        l.oracle_brute.argtypes = [_P, _P, ctypes.c_uint64, _P, _P, ctypes.c_int]
        l.oracle_ray_setup.argtypes = [_P, _P, _P]
        l.oracle_tri_simd.restype = ctypes.c_int
        l.oracle_tri_simd.argtypes = [_P, _P, _P]
        l.oracle_tri_scalar.restype = ctypes.c_int
        l.oracle_tri_scalar.argtypes = [_P, _P, _P]
        l.oracle_bbox8.restype = ctypes.c_int
        l.oracle_bbox8.argtypes = [_P, _P, _P, _P]
        l.oracle_root_slab.restype = ctypes.c_float
        l.oracle_root_slab.argtypes = [_P, _P]
        l.oracle_quad.restype = ctypes.c_int
        l.oracle_quad.argtypes = [_P, _P, _P]
        l.oracle_simd_rcp8.argtypes = [_P, _P]
        l.oracle_simd_minreduce8.restype = ctypes.c_float
        l.oracle_simd_minreduce8.argtypes = [_P]
        l.oracle_simd_movemask8.restype = ctypes.c_int
        l.oracle_simd_movemask8.argtypes = [_P]
        _lib = l
    return _lib


def _ptr(a):
    return a.ctypes.data if a is not None and a.size else None


def quads_array(quads) -> np.ndarray:
    if not quads:
        return np.zeros((0, 34), np.float32)
    rows = []
    for q in quads:
        rows.append(np.concatenate([np.asarray(q["mat"], np.float32).reshape(16),
                                    np.asarray(q["inv"], np.float32).reshape(16),
                                    np.array([q["size_x"], q["size_y"]], np.float32)]))
    return np.ascontiguousarray(np.stack(rows).astype(np.float32))


class OracleBvh:
    """The restated Fast BVH of SORT (width 8 = Obvh/AVX, 4 = Qbvh/SSE)."""

    def __init__(self, pos, idx, quads=None, width=8, max_depth=28, max_leaf=16):
        self.pos = np.ascontiguousarray(np.asarray(pos, np.float32).reshape(-1, 3))
        self.idx = np.ascontiguousarray(np.asarray(idx, np.uint32).reshape(-1, 3))
        self.q = quads_array(quads)
        self.h = lib().oracle_build(_ptr(self.pos), len(self.pos), _ptr(self.idx), len(self.idx),
                                    _ptr(self.q), len(self.q), width, max_depth, max_leaf)
        self.width = width

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def info(self):
        out = np.zeros(5, np.uint64)
        lib().oracle_info(self.h, out.ctypes.data)
        return dict(zip(["depth", "nodes", "leaves", "prims", "valid"], [int(x) for x in out]))

    def bbox(self):
        out = np.zeros(6, np.float32)
        lib().oracle_scene_bbox(self.h, out.ctypes.data)
        return out

    def prim_order(self):
        n = self.info()["prims"]
        out = np.zeros(n, np.uint32)
        if n:
            lib().oracle_prim_order(self.h, out.ctypes.data)
        return out

    def tree_dump(self):
        n = lib().oracle_tree_dump(self.h, None, 0)
        out = np.zeros(n, np.uint32)
        if n:
            lib().oracle_tree_dump(self.h, out.ctypes.data, n)
        return out

    def closest(self, rays, threads=1, stats=None):
        rays = np.ascontiguousarray(np.asarray(rays, np.float32).reshape(-1, 8))
        hits = np.zeros((len(rays), 4), np.float32)
        st = np.zeros(5, np.uint64)
        secs = lib().oracle_trace(self.h, _ptr(rays), len(rays), _ptr(hits), None, 0, threads, st.ctypes.data)
        if stats is not None:
            stats.update(dict(zip(["rays", "inner_nodes", "leaves", "tris", "quads"], [int(x) for x in st])))
            stats["seconds"] = secs
        return hits

    def occluded(self, rays, threads=1, stats=None):
        rays = np.ascontiguousarray(np.asarray(rays, np.float32).reshape(-1, 8))
        occ = np.zeros(len(rays), np.uint8)
        st = np.zeros(5, np.uint64)
        secs = lib().oracle_trace(self.h, _ptr(rays), len(rays), None, _ptr(occ), 1, threads, st.ctypes.data)
        if stats is not None:
            stats.update(dict(zip(["rays", "inner_nodes", "leaves", "tris", "quads"], [int(x) for x in st])))
            stats["seconds"] = secs
        return occ

    def brute_closest(self, rays):
        rays = np.ascontiguousarray(np.asarray(rays, np.float32).reshape(-1, 8))
        hits = np.zeros((len(rays), 4), np.float32)
        lib().oracle_brute(self.h, _ptr(rays), len(rays), _ptr(hits), None, 0)
        return hits

    def brute_occluded(self, rays):
        rays = np.ascontiguousarray(np.asarray(rays, np.float32).reshape(-1, 8))
        occ = np.zeros(len(rays), np.uint8)
        lib().oracle_brute(self.h, _ptr(rays), len(rays), None, _ptr(occ), 1)
        return occ


def ray_setup(ray):
    r = np.ascontiguousarray(np.asarray(ray, np.float32).reshape(8))
    k = np.zeros(3, np.int32)
    f = np.zeros(9, np.float32)
    lib().oracle_ray_setup(r.ctypes.data, k.ctypes.data, f.ctypes.data)
    return k, f


def tri_simd(ray, tri9):
    r = np.ascontiguousarray(np.asarray(ray, np.float32).reshape(8))
    t = np.ascontiguousarray(np.asarray(tri9, np.float32).reshape(9))
    out = np.zeros(3, np.float32)
    ok = lib().oracle_tri_simd(r.ctypes.data, t.ctypes.data, out.ctypes.data)
    return (bool(ok), out)


def tri_scalar(ray, tri9):
    r = np.ascontiguousarray(np.asarray(ray, np.float32).reshape(8))
    t = np.ascontiguousarray(np.asarray(tri9, np.float32).reshape(9))
    out = np.zeros(3, np.float32)
    ok = lib().oracle_tri_scalar(r.ctypes.data, t.ctypes.data, out.ctypes.data)
    return (bool(ok), out)


def bbox8(ray, mn, mx, valid):
    """mn, mx: (3,8) planes; valid: 8 flags. Returns (mask, fmin[8])."""
    r = np.ascontiguousarray(np.asarray(ray, np.float32).reshape(8))
    boxes = np.ascontiguousarray(np.concatenate([np.asarray(mn, np.float32).reshape(24),
                                                 np.asarray(mx, np.float32).reshape(24)]))
    v = np.ascontiguousarray(np.asarray(valid, np.int32).reshape(8))
    f = np.zeros(8, np.float32)
    m = lib().oracle_bbox8(r.ctypes.data, boxes.ctypes.data, v.ctypes.data, f.ctypes.data)
    return m, f


def root_slab(ray, box6):
    r = np.ascontiguousarray(np.asarray(ray, np.float32).reshape(8))
    b = np.ascontiguousarray(np.asarray(box6, np.float32).reshape(6))
    return float(lib().oracle_root_slab(r.ctypes.data, b.ctypes.data))


def quad(ray, q):
    r = np.ascontiguousarray(np.asarray(ray, np.float32).reshape(8))
    qa = quads_array([q])[0]
    t = np.zeros(1, np.float32)
    ok = lib().oracle_quad(r.ctypes.data, qa.ctypes.data, t.ctypes.data)
    return bool(ok), float(t[0])


def simd_rcp8(x):
    a = np.ascontiguousarray(np.asarray(x, np.float32).reshape(8))
    out = np.zeros(8, np.float32)
    lib().oracle_simd_rcp8(a.ctypes.data, out.ctypes.data)
    return out


def simd_minreduce8(x):
    a = np.ascontiguousarray(np.asarray(x, np.float32).reshape(8))
    return float(lib().oracle_simd_minreduce8(a.ctypes.data))


def simd_movemask8(x):
    a = np.ascontiguousarray(np.asarray(x, np.float32).reshape(8))
    return int(lib().oracle_simd_movemask8(a.ctypes.data))
