"""Seeded synthetic workloads (numpy, float32) for the intersection and path-tracing benchmarks.

Everything is generated with ``numpy.random.Generator(PCG64(seed))`` so the same seed gives the
same bytes on every host with this numpy version. Seeds are fixed constants below.

* ``cornell_sphere``: room [0,1]^3 (6 walls) + a 6-face pedestal = 12 axis-aligned quads
  (24 triangles), a 64x32-segment UV sphere (4,032 triangles) and one 0.25 x 0.25 area-light
  quad just below the ceiling facing -y.
* ``soup``: n triangles, centroids U[0,1]^3, size s ~ LogNormal(ln 0.004, 0.5), vertices
  c + s * (a_k - mean(a)) with a_k uniform on the unit sphere.
* ``pinhole_rays`` / ``incoherent_rays`` / ``shadow_rays``: the coherent, incoherent and shadow
  ray sets of the 1M-triangle intersection workload.
"""
from __future__ import annotations

import math

import numpy as np

FLT_MAX = np.float32(3.4028234663852886e38)

SEED_CORNELL = 0x5EED0001
SEED_SOUP = 0x5EED0002
SEED_PT = 0x5EED0003
SEED_BATCH = 0x5EED0004
SEED_SERVE = 0x5EED0005


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


# ---------------------------------------------------------------------------------------------
# geometry

def _quad_tris(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3) of a planar quad."""
    return [np.array([p0, p1, p2], np.float32), np.array([p0, p2, p3], np.float32)]


def axis_box_faces(lo, hi):
    """Six outward-facing quads of an axis-aligned box, as 12 triangles."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    faces = [
        ((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1)),  # -y
        ((x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0)),  # +y
        ((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0)),  # -x
        ((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1)),  # +x
        ((x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0)),  # -z
        ((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)),  # +z
    ]
    tris = []
    for f in faces:
        tris += _quad_tris(*f)
    return tris


def uv_sphere(center, radius, segments=64, rings=32):
    """UV sphere: a segments x rings grid of quads split into 2 triangles each (4,096 for
    64 x 32). The 64 zero-area triangles that touch the north pole are dropped; the 64 at the
    south pole are kept (they exercise the det == 0 rejection), giving 4,032 triangles."""
    cx, cy, cz = center
    tris = []
    def vert(i, j):
        th = math.pi * i / rings
        ph = 2.0 * math.pi * (j % segments) / segments
        return (cx + radius * math.sin(th) * math.cos(ph), cy + radius * math.cos(th),
                cz + radius * math.sin(th) * math.sin(ph))
    for i in range(rings):
        for j in range(segments):
            a, b, c, d = vert(i, j), vert(i, j + 1), vert(i + 1, j + 1), vert(i + 1, j)
            t1 = np.array([a, b, c], np.float32)
            t2 = np.array([a, c, d], np.float32)
            if i == 0:
                tris.append(t2)  # (a, b) coincide at the north pole: t1 has zero area
            else:
                tris.append(t1)
                tris.append(t2)
    return tris


def translate_quad(center, size, facing_down=True):
    """Light-to-world transform of a size x size quad centred at `center`, normal -y when
    facing_down (a 180-degree rotation about x), and its inverse. Row-major 4x4."""
    cx, cy, cz = center
    if facing_down:
        mat = np.array([[1, 0, 0, cx], [0, -1, 0, cy], [0, 0, -1, cz], [0, 0, 0, 1]], np.float32)
        inv = np.array([[1, 0, 0, -cx], [0, -1, 0, cy], [0, 0, -1, cz], [0, 0, 0, 1]], np.float32)
    else:
        mat = np.array([[1, 0, 0, cx], [0, 1, 0, cy], [0, 0, 1, cz], [0, 0, 0, 1]], np.float32)
        inv = np.array([[1, 0, 0, -cx], [0, 1, 0, -cy], [0, 0, 1, -cz], [0, 0, 0, 1]], np.float32)
    return {"mat": mat, "inv": inv, "size_x": size, "size_y": size}


def tris_to_mesh(tris):
    t = np.stack(tris).astype(np.float32)  # (T,3,3)
    pos = t.reshape(-1, 3)
    idx = np.arange(pos.shape[0], dtype=np.uint32).reshape(-1, 3)
    return np.ascontiguousarray(pos), np.ascontiguousarray(idx)


def cornell_sphere():
    """Returns dict(pos, idx, quads, mat_id (per triangle), groups) for the Cornell + sphere
    scene. Triangle order: room walls (12), pedestal (12), sphere (4,032)."""
    room = axis_box_faces((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    # room faces must face inward: flip the winding of every room triangle
    room = [t[[0, 2, 1]] for t in room]
    pedestal = axis_box_faces((0.35, 0.0, 0.35), (0.65, 0.2, 0.65))
    sphere = uv_sphere((0.5, 0.5, 0.5), 0.25)
    tris = room + pedestal + sphere
    pos, idx = tris_to_mesh(tris)
    # materials: 0 white wall, 1 red (-x wall), 2 green (+x wall), 3 pedestal, 4 sphere
    mat = np.zeros(len(tris), np.uint32)
    mat[4:6] = 1   # -x wall
    mat[6:8] = 2   # +x wall
    mat[12:24] = 3
    mat[24:] = 4
    light = translate_quad((0.5, 0.999, 0.5), 0.25, facing_down=True)
    return {"pos": pos, "idx": idx, "quads": [light], "mat": mat,
            "ntris": len(tris), "counts": {"room": 12, "pedestal": 12, "sphere": len(sphere)}}


def soup(n_tris: int = 1 << 20, seed: int = SEED_SOUP):
    """Seeded triangle soup (positions (3n,3), indices (n,3))."""
    g = rng(seed)
    c = g.random((n_tris, 3), dtype=np.float32)
    s = np.exp(g.normal(np.log(0.004), 0.5, n_tris)).astype(np.float32)
    a = g.normal(size=(n_tris, 3, 3)).astype(np.float32)
    a /= np.linalg.norm(a, axis=2, keepdims=True).astype(np.float32)
    a -= a.mean(axis=1, keepdims=True)
    v = (c[:, None, :] + s[:, None, None] * a).astype(np.float32)
    pos = np.ascontiguousarray(v.reshape(-1, 3))
    idx = np.arange(3 * n_tris, dtype=np.uint32).reshape(-1, 3)
    return pos, idx


# ---------------------------------------------------------------------------------------------
# rays: (n, 8) float32 [ox, oy, oz, tmin, dx, dy, dz, tmax]

def make_rays(o, d, tmin, tmax):
    n = o.shape[0]
    r = np.empty((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def pinhole_rays(width=2048, height=2048, eye=(0.5, 0.5, -1.5), target=(0.5, 0.5, 0.5),
                 up=(0.0, 1.0, 0.0), vfov_deg=40.0):
    """Coherent camera rays through pixel centres, row-major, tmin 0, tmax FLT_MAX."""
    eye = np.array(eye, np.float32)
    fwd = np.array(target, np.float32) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.array(up, np.float32)).astype(np.float32)
    right /= np.linalg.norm(right)
    upv = np.cross(right, fwd).astype(np.float32)
    th = np.float32(math.tan(math.radians(vfov_deg) * 0.5))
    aspect = np.float32(width / height)
    xs = ((np.arange(width, dtype=np.float32) + 0.5) / width * 2.0 - 1.0) * th * aspect
    ys = (1.0 - (np.arange(height, dtype=np.float32) + 0.5) / height * 2.0) * th
    X, Y = np.meshgrid(xs, ys)
    d = fwd[None, :] + X.reshape(-1, 1) * right[None, :] + Y.reshape(-1, 1) * upv[None, :]
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    o = np.broadcast_to(eye, d.shape).astype(np.float32)
    return make_rays(o, d, 0.0, FLT_MAX)


def _tri_normals(pos, idx, prims):
    v0 = pos[idx[prims, 0]]
    v1 = pos[idx[prims, 1]]
    v2 = pos[idx[prims, 2]]
    n = np.cross(v2 - v0, v1 - v0).astype(np.float32)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    ln[ln == 0] = 1.0
    return (n / ln).astype(np.float32)


def _onb(n):
    """Orthonormal tangent frames (t, b) for unit normals n (Duff et al. 2017)."""
    s = np.where(n[:, 2] >= 0, 1.0, -1.0).astype(np.float32)
    a = -1.0 / (s + n[:, 2])
    b = n[:, 0] * n[:, 1] * a
    t = np.stack([1.0 + s * n[:, 0] * n[:, 0] * a, s * b, -s * n[:, 0]], 1).astype(np.float32)
    bb = np.stack([b, s + n[:, 1] * n[:, 1] * a, -n[:, 1]], 1).astype(np.float32)
    return t, bb


def hit_origins(cam_rays, hits, pos, idx, count, seed=SEED_SOUP + 1):
    """Origins on random camera hits (hit i = random index over the rays that hit) with
    geometric normals facing the incoming ray."""
    prims = hits[:, 1].view(np.uint32)
    ok = np.nonzero(prims != 0xFFFFFFFF)[0]
    if ok.size == 0:
        raise ValueError("no camera ray hit the scene")
    g = rng(seed)
    pick = ok[g.integers(0, ok.size, count)]
    o = cam_rays[pick, 0:3]
    d = cam_rays[pick, 4:7]
    t = hits[pick, 0:1]
    p = (o + t * d).astype(np.float32)
    n = _tri_normals(pos, idx, prims[pick])
    flip = np.sum(n * d, axis=1) > 0
    n[flip] = -n[flip]
    return p, n, g


def incoherent_rays(cam_rays, hits, pos, idx, count=1 << 22, seed=SEED_SOUP + 1):
    """Cosine-hemisphere bounce rays from random camera hits; tmin 1e-4 (the path tracer's
    bounce epsilon), tmax FLT_MAX."""
    p, n, g = hit_origins(cam_rays, hits, pos, idx, count, seed)
    u1 = g.random(count, dtype=np.float32)
    u2 = g.random(count, dtype=np.float32)
    r = np.sqrt(u1)
    phi = np.float32(2.0 * math.pi) * u2
    lx, lz, ly = r * np.cos(phi), r * np.sin(phi), np.sqrt(np.maximum(0.0, 1.0 - u1))
    t, b = _onb(n)
    d = (t * lx[:, None] + n * ly[:, None] + b * lz[:, None]).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return make_rays(p, d.astype(np.float32), np.float32(1e-4), FLT_MAX)


def shadow_rays(cam_rays, hits, pos, idx, count=1 << 22, seed=SEED_SOUP + 2,
                light_center=(0.5, 1.2, 0.5), light_size=0.25):
    """Segments from random camera hits to uniform points on a light quad at y = 1.2:
    tmin 0.01, tmax = length - 0.01 (the area-light shadow ray, light/area.cpp:48-50)."""
    p, _, g = hit_origins(cam_rays, hits, pos, idx, count, seed)
    u = g.random((count, 2), dtype=np.float32)
    lc = np.array(light_center, np.float32)
    lp = np.stack([lc[0] + (u[:, 0] - 0.5) * light_size, np.full(count, lc[1], np.float32),
                   lc[2] + (u[:, 1] - 0.5) * light_size], 1).astype(np.float32)
    v = lp - p
    ln = np.linalg.norm(v, axis=1).astype(np.float32)
    d = (v / ln[:, None]).astype(np.float32)
    return make_rays(p, d, np.float32(0.01), (ln - np.float32(0.01)).astype(np.float32))


def random_rays(count, seed, lo=-0.5, hi=1.5, tmin=0.0, tmax=FLT_MAX):
    """Uniform origins in [lo,hi]^3 with uniform directions (for small parity cases)."""
    g = rng(seed)
    o = g.uniform(lo, hi, (count, 3)).astype(np.float32)
    d = g.normal(size=(count, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return make_rays(o, d.astype(np.float32), np.float32(tmin), np.float32(tmax))
