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
        l.oracle_attenuation.argtypes = [_P, _P, ctypes.c_uint64, _P, _P, ctypes.c_uint32, ctypes.c_int]
        l.oracle_multi.argtypes = [_P, _P, ctypes.c_uint64, _P, _P, _P, _P, ctypes.c_int]
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

    def attenuation(self, rays, transparency, threads=1, max_steps=4096):
        """Transparent-shadow attenuation (Scene::GetAttenuation): (n,4) [r, g, b, surfaces passed].
        transparency: (n_prims, 4) per primitive id [r, g, b, 1 if the material has transparency]."""
        rays = np.ascontiguousarray(np.asarray(rays, np.float32).reshape(-1, 8))
        tr = np.ascontiguousarray(np.asarray(transparency, np.float32).reshape(-1, 4))
        out = np.zeros((len(rays), 4), np.float32)
        lib().oracle_attenuation(self.h, _ptr(rays), len(rays), _ptr(tr), _ptr(out), max_steps, threads)
        return out

    def multi(self, rays, ray_material, tri_material, threads=1):
        """BSSRDF multi-hit query (Fbvh::GetIntersect with BSSRDFIntersections): ((n,4,4) records
        [t, prim bits, u, v], (n,) counts)."""
        rays = np.ascontiguousarray(np.asarray(rays, np.float32).reshape(-1, 8))
        rm = np.ascontiguousarray(np.asarray(ray_material, np.uint32).reshape(-1))
        tm = np.ascontiguousarray(np.asarray(tri_material, np.uint32).reshape(-1))
        hits = np.zeros((len(rays), 4, 4), np.float32)
        cnt = np.zeros(len(rays), np.uint32)
        lib().oracle_multi(self.h, _ptr(rays), len(rays), _ptr(rm), _ptr(tm), _ptr(hits), _ptr(cnt), threads)
        return hits, cnt

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


# ---- path tracer (pt_oracle.cpp) ----------------------------------------------------------------

def _pt_sigs():
    l = lib()
    if getattr(l, "_pt_ready", False):
        return l
    l.oracle_pt_create.restype = _P
    l.oracle_pt_create.argtypes = [_P, _P, ctypes.c_uint32, _P, ctypes.c_uint32, _P, _P, _P, _P, ctypes.c_uint32,
                                   _P, ctypes.c_uint32, _P, _P, _P]
    l.oracle_pt_destroy.argtypes = [_P]
    l.oracle_pt_set_clamp.argtypes = [_P, ctypes.c_float]
    l.oracle_pt_render.restype = ctypes.c_double
    l.oracle_pt_render.argtypes = [_P, _P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, _P, _P]
    l.oracle_pt_li.argtypes = [_P, _P, _P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, _P]
    l.oracle_pt_dump.restype = ctypes.c_double
    l.oracle_pt_dump.argtypes = [_P, _P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, _P, _P, _P]
    l.oracle_bxdf.argtypes = [_P, ctypes.c_int, ctypes.c_uint32, _P, _P, _P, _P]
    l._pt_ready = True
    return l


MAT_FIELDS = ["metallic", "specular", "specular_tint", "sheen", "sheen_tint", "clearcoat", "clearcoat_gloss"]


def material_record(m) -> np.ndarray:
    """22-word record: type, base[3], trans[3], roughness, ior, eta[3], k[3], 7 Disney scalars."""
    rec = np.zeros(22, np.float32)
    rec[1:4] = m["base"]
    rec[4:7] = m["trans"]
    rec[7] = m["roughness"]
    rec[8] = m["ior"]
    rec[9:12] = m["eta"]
    rec[12:15] = m["k"]
    rec[15:22] = [m[f] for f in MAT_FIELDS]
    rec.view(np.uint32)[0] = m["type"]
    return rec


class OraclePT:
    """The restated SORT path tracer over an OracleBvh (same geometry, quads first-class)."""

    def __init__(self, scene, width=8, max_depth=28, max_leaf=16):
        l = _pt_sigs()
        self.bvh = OracleBvh(scene["pos"], scene["idx"], scene["quads"], width=width, max_depth=max_depth,
                             max_leaf=max_leaf)
        self._keep = []

        def arr(a, dt):
            if a is None:
                return None
            x = np.ascontiguousarray(np.asarray(a, dt))
            self._keep.append(x)
            return x

        pos = arr(scene["pos"], np.float32)
        idx = arr(scene["idx"], np.uint32)
        nrm = arr(scene.get("normals"), np.float32)
        tan = arr(scene.get("tangents"), np.float32)
        tm = arr(scene.get("tri_material"), np.uint32)
        mats = arr(np.stack([material_record(m) for m in scene["materials"]]), np.float32)
        q34 = arr(quads_array(scene["quads"]), np.float32)
        qr = arr(scene.get("quad_radiance", np.zeros((len(scene["quads"]), 3))), np.float32)
        env = arr(scene.get("env", (0, 0, 0)), np.float32)
        bb = arr(self.bvh.bbox(), np.float32)
        self.h = l.oracle_pt_create(self.bvh.h, _ptr(pos), len(pos), _ptr(idx), len(idx), _ptr(nrm), _ptr(tan),
                                    _ptr(tm), _ptr(mats), len(mats), _ptr(q34), len(q34), _ptr(qr), _ptr(env),
                                    _ptr(bb))
        if scene.get("clamp", 0.0):
            l.oracle_pt_set_clamp(self.h, float(scene["clamp"]))

    def __del__(self):
        if getattr(self, "h", None):
            _pt_sigs().oracle_pt_destroy(self.h)
            self.h = None

    def render(self, cam, width, height, spp, max_depth, seed, threads=8, crop=None, stats=None):
        """cam: dict(eye, target, up, fov_y_deg). crop = (pix0, pix1) linear pixel range.
        Returns (npix, 3) radiance (mean of valid samples)."""
        import math
        c = np.array([*cam["eye"], *cam["target"], *cam["up"], math.radians(cam["fov_y_deg"])], np.float32)
        p0, p1 = crop if crop else (0, width * height)
        out = np.zeros((p1 - p0, 3), np.float32)
        st = np.zeros(2, np.uint64)
        secs = _pt_sigs().oracle_pt_render(self.h, c.ctypes.data, width, height, p0, p1, spp, max_depth, seed,
                                           threads, out.ctypes.data, st.ctypes.data)
        if stats is not None:
            stats.update(extension_rays=int(st[0]), shadow_rays=int(st[1]), seconds=secs)
        return out

    def ray_dump(self, cam, width, height, spp, max_depth, seed, pix0, pix1, threads=8):
        """Every extension ray (bounce 0 = camera ray) of the paths of pixels [pix0, pix1) with its
        closest hit, at slot ((pixel - pix0) * spp + sample) * max_depth + bounce. Returns
        (rays (n, 8) float32, hits (n, 4) float32 with prim as uint32 bits, valid (n,) bool)."""
        import math
        c = np.array([*cam["eye"], *cam["target"], *cam["up"], math.radians(cam["fov_y_deg"])], np.float32)
        n = (pix1 - pix0) * spp * max_depth
        rays = np.zeros((n, 8), np.float32)
        hits = np.zeros((n, 4), np.float32)
        valid = np.zeros(n, np.uint8)
        _pt_sigs().oracle_pt_dump(self.h, c.ctypes.data, width, height, pix0, pix1, spp, max_depth, seed, threads,
                                  rays.ctypes.data, hits.ctypes.data, valid.ctypes.data)
        return rays, hits, valid.astype(bool)

    def li(self, rays, keys, max_depth, seed, threads=8):
        rays = np.ascontiguousarray(np.asarray(rays, np.float32).reshape(-1, 8))
        keys = np.ascontiguousarray(np.asarray(keys, np.uint32).reshape(-1, 2))
        out = np.zeros((len(rays), 3), np.float32)
        _pt_sigs().oracle_pt_li(self.h, rays.ctypes.data, keys.ctypes.data, len(rays), max_depth, seed, threads,
                                out.ctypes.data)
        return out


def _bxdf(mat, mode, wo, wi=None, u=None):
    rec = None if mat is None else material_record(mat)
    wo = np.ascontiguousarray(np.asarray(wo, np.float32).reshape(-1, 3))
    n = len(wo)
    wi = None if wi is None else np.ascontiguousarray(np.asarray(wi, np.float32).reshape(-1, 3))
    u = None if u is None else np.ascontiguousarray(np.asarray(u, np.float32).reshape(-1, 7))
    out = np.zeros((n, 7), np.float32)
    _pt_sigs().oracle_bxdf(_ptr(rec), mode, n, _ptr(wo), _ptr(wi), _ptr(u), _ptr(out))
    return out


def bxdf_f(mat, wo, wi):
    """f(wo, wi) * |cos wi| (SORT's convention) for arrays of local directions: (n, 3)."""
    return _bxdf(mat, 0, wo, wi=wi)[:, 0:3]


def bxdf_pdf(mat, wo, wi):
    return _bxdf(mat, 1, wo, wi=wi)[:, 0]


def bxdf_sample(mat, wo, u7):
    """Returns f (n,3), wi (n,3), pdf (n,)."""
    out = _bxdf(mat, 2, wo, u=u7)
    return out[:, 0:3], out[:, 3:6], out[:, 6]
