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


def face_normals_per_vertex(pos, idx):
    """Per-vertex normals equal to each (unshared-vertex) triangle's CCW face normal."""
    p0, p1, p2 = pos[idx[:, 0]], pos[idx[:, 1]], pos[idx[:, 2]]
    n = np.cross(p1 - p0, p2 - p0).astype(np.float32)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    ln[ln == 0] = 1.0
    n = (n / ln).astype(np.float32)
    out = np.zeros_like(pos)
    out[idx[:, 0]] = n
    out[idx[:, 1]] = n
    out[idx[:, 2]] = n
    return out


def cornell_sphere(light_radiance=(17.0, 17.0, 17.0), sphere_material=None):
    """Closed room [0,1]^3 (6 walls, inward normals) + a 6-face pedestal = 12 axis-aligned quads
    (24 triangles), a 64x32 UV sphere (4,032 triangles, radius 0.25 at the centre, smooth
    normals) and one 0.25 x 0.25 area light at y = 0.999 facing -y.

    Triangle order: room walls (12), pedestal (12), sphere (4,032). Materials: 0 white Lambert
    0.7, 1 red (0.7, 0.2, 0.2) on the -x wall, 2 green (0.2, 0.7, 0.2) on the +x wall, 3 the
    pedestal (white 0.7), 4 the sphere (white Lambert 0.7 unless `sphere_material` is given).
    The camera sits inside the room by the z = 0 wall looking at +z."""
    from .integrator import lambert
    room = axis_box_faces((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    room = [t[[0, 2, 1]] for t in room]  # inward-facing walls
    pedestal = axis_box_faces((0.35, 0.0, 0.35), (0.65, 0.2, 0.65))
    sphere = uv_sphere((0.5, 0.5, 0.5), 0.25)
    tris = room + pedestal + sphere
    pos, idx = tris_to_mesh(tris)
    nrm = face_normals_per_vertex(pos, idx)
    sv = idx[24:].reshape(-1)
    d = pos[sv] - np.array([0.5, 0.5, 0.5], np.float32)
    nrm[sv] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    mat = np.zeros(len(tris), np.uint32)
    mat[4:6] = 1   # -x wall
    mat[6:8] = 2   # +x wall
    mat[12:24] = 3
    mat[24:] = 4
    materials = [lambert((0.7, 0.7, 0.7)), lambert((0.7, 0.2, 0.2)), lambert((0.2, 0.7, 0.2)),
                 lambert((0.7, 0.7, 0.7)), sphere_material or lambert((0.7, 0.7, 0.7))]
    light = translate_quad((0.5, 0.999, 0.5), 0.25, facing_down=True)
    return {"pos": pos, "idx": idx, "quads": [light], "normals": nrm, "tri_material": mat,
            "materials": materials, "quad_radiance": np.array([light_radiance], np.float32),
            "env": (0.0, 0.0, 0.0), "ntris": len(tris),
            "camera": dict(eye=(0.5, 0.5, 0.02), target=(0.5, 0.5, 1.0), up=(0.0, 1.0, 0.0), fov_y_deg=70.0),
            "counts": {"room": 12, "pedestal": 12, "sphere": len(sphere)}}


def icosphere(level=5):
    """Unit icosphere: the icosahedron (level 1) subdivided level-1 times: 20 * 4^(level-1)
    triangles (5,120 at level 5), shared vertices, outward CCW winding."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(x, np.float64) / np.linalg.norm(x) for x in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level - 1):
        cache = {}
        nf = []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts, np.float32), np.array(faces, np.uint32)


def spheres_scene(n_spheres=400, level=5, seed=SEED_PT):
    """Ground quad + n seeded icospheres (radii U[0.05, 0.4], resting on the ground) baked into
    one mesh; materials drawn 50% Lambert / 30% GGX (roughness U[0.05, 0.8]) / 20% dielectric
    (IOR 1.5); a constant environment and 4 quad lights facing down."""
    from .integrator import dielectric, ggx, lambert
    g = rng(seed)
    uv, uf = icosphere(level)
    nv, nf = len(uv), len(uf)
    G = 6.0
    ground_p = np.array([[-G, 0, -G], [G, 0, -G], [G, 0, G], [-G, 0, G]], np.float32)
    ground_i = np.array([[0, 2, 1], [0, 3, 2]], np.uint32)  # CCW seen from +y
    radii = g.uniform(0.05, 0.4, n_spheres).astype(np.float32)
    cx = g.uniform(-4.5, 4.5, n_spheres).astype(np.float32)
    cz = g.uniform(-4.5, 4.5, n_spheres).astype(np.float32)
    pos = np.empty((4 + n_spheres * nv, 3), np.float32)
    nrm = np.empty_like(pos)
    idx = np.empty((2 + n_spheres * nf, 3), np.uint32)
    pos[:4] = ground_p
    nrm[:4] = (0, 1, 0)
    idx[:2] = ground_i
    for i in range(n_spheres):
        c = np.array([cx[i], radii[i], cz[i]], np.float32)
        pos[4 + i * nv: 4 + (i + 1) * nv] = c + radii[i] * uv
        nrm[4 + i * nv: 4 + (i + 1) * nv] = uv
        idx[2 + i * nf: 2 + (i + 1) * nf] = uf + np.uint32(4 + i * nv)
    kind = g.random(n_spheres)
    rough = g.uniform(0.05, 0.8, n_spheres).astype(np.float32)
    albedo = g.uniform(0.2, 0.9, (n_spheres, 3)).astype(np.float32)
    materials = [lambert((0.6, 0.6, 0.6))]
    sphere_mat = np.empty(n_spheres, np.uint32)
    for i in range(n_spheres):
        sphere_mat[i] = len(materials)
        if kind[i] < 0.5:
            materials.append(lambert(tuple(albedo[i])))
        elif kind[i] < 0.8:
            materials.append(ggx(tuple(albedo[i]), float(rough[i])))
        else:
            materials.append(dielectric((1, 1, 1), (1, 1, 1), roughness=float(rough[i]) * 0.25, ior=1.5))
    tri_mat = np.zeros(len(idx), np.uint32)
    tri_mat[2:] = np.repeat(sphere_mat, nf)
    lights = [translate_quad((x, 3.5, z), 1.0, facing_down=True) for x, z in [(-2.5, -2.5), (2.5, -2.5), (-2.5, 2.5), (2.5, 2.5)]]
    return {"pos": pos, "idx": idx, "quads": lights, "normals": nrm, "tri_material": tri_mat,
            "materials": materials, "quad_radiance": np.full((4, 3), 6.0, np.float32), "env": (0.25, 0.3, 0.35),
            "ntris": len(idx),
            "camera": dict(eye=(0.0, 3.0, 9.0), target=(0.0, 0.5, 0.0), up=(0.0, 1.0, 0.0), fov_y_deg=45.0)}


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
