"""Pins the oracle's plain queries (closest hit, any hit) by brute force over every primitive.

The brute force here is an independent numpy transcription of the per-primitive arithmetic:
`Ray::Prepare` (math/ray.h:60-69), the watertight shear test `intersectTriangleInner_SIMD`
(simd/simd_triangle.h:212-300: p -= o, permute, p.x += p.y*sx and p.z += p.y*sz as multiply then
add, edge functions, all >= 0 or all <= 0, det = e0 + (e1 + e2) with an ordered det != 0,
t = ((e0*p0y + e1*p1y) + e2*p2y) * (1/det), window (tmin, tmax]) and `Quad::GetIntersect`
(shape/quad.cpp:66-94: ray into light space with Matrix::TransformPoint/TransformVector,
t = -o.y/d.y, reject t > limit, t <= tmin, t > tmax, then the half-size checks). numpy float32
arithmetic rounds every operation and never fuses, like the reference's AVX build without FMA.

What the tree walk must give (`Fbvh::GetIntersect`, accel/fast_bvh.hpp:305-504;
`intersectTriangle_SIMD`, simd_triangle.h:350-383; shadow mode fast_bvh.hpp:362-373 and
accelerator.cpp:26-39): the smallest accepted t over all primitives, bitwise, because each
primitive's t is computed by the same float operations whatever the order of visits; the same
primitive unless another primitive has exactly that t (then the visit order decides); and u, v
bitwise whenever the primitive is the same. Any hit: true exactly when some primitive is accepted
in (tmin, tmax]. Reference quirk kept by both: the root test (math/bbox.h:124-156) starts from the
ray's tmin, so a ray whose window starts below zero with its origin inside the scene box gets a
negative entry distance and `GetIntersect` returns no hit at all.
"""
import numpy as np
import pytest

from oracle.oracle import OracleBvh
from sort_amd import scenes

FLT_MAX = np.float32(3.4028234663852886e38)
F = np.float32


# ---- independent brute force ------------------------------------------------------------------

def np_prepare(d):
    """Ray::Prepare (math/ray.h:26-30, 60-69): major axis with ties to the higher index."""
    ad = np.abs(d)
    if ad[0] > ad[1] and ad[0] > ad[2]:
        ky = 0
    else:
        ky = 1 if ad[1] > ad[2] else 2
    kz = (ky + 1) % 3
    kx = (kz + 1) % 3
    with np.errstate(all="ignore"):
        sx = F(-d[kx]) / d[ky]
        sz = F(-d[kz]) / d[ky]
        sy = F(1.0) / d[ky]
    return kx, ky, kz, F(sx), F(sy), F(sz)


def np_triangles(ray, tri):
    """Every triangle against one ray: (accepted mask, t, u, v) per triangle. tri: (n, 3, 3)."""
    o = ray[0:3]
    d = ray[4:7]
    tmin, tmax = ray[3], ray[7]
    kx, ky, kz, sx, sy, sz = np_prepare(d)
    with np.errstate(all="ignore"):
        p = (tri - o[None, None, :]).astype(F)                     # (n, 3 verts, 3 axes)
        px, py, pz = p[:, :, kx], p[:, :, ky], p[:, :, kz]
        px = (py * sx + px).astype(F)
        pz = (py * sz + pz).astype(F)
        e0 = (px[:, 1] * pz[:, 2] - pz[:, 1] * px[:, 2]).astype(F)
        e1 = (px[:, 2] * pz[:, 0] - pz[:, 2] * px[:, 0]).astype(F)
        e2 = (px[:, 0] * pz[:, 1] - pz[:, 0] * px[:, 1]).astype(F)
        c0 = (e0 >= 0) & (e1 >= 0) & (e2 >= 0)
        c1 = (e0 <= 0) & (e1 <= 0) & (e2 <= 0)
        det = (e0 + (e1 + e2)).astype(F)
        ok = (c0 | c1) & ((det < 0) | (det > 0))                   # ordered det != 0
        rcp = (F(1.0) / det).astype(F)
        py = (py * sy).astype(F)
        t = (e0 * py[:, 0]).astype(F)
        t = (e1 * py[:, 1] + t).astype(F)
        t = (e2 * py[:, 2] + t).astype(F)
        t = (t * rcp).astype(F)
        ok &= (t > tmin) & (t <= tmax)
        u = (e1 * rcp).astype(F)
        v = (e2 * rcp).astype(F)
    return ok, t, u, v


def np_transform_point(m, p):
    x = p[0] * m[0] + p[1] * m[1] + p[2] * m[2] + m[3]
    y = p[0] * m[4] + p[1] * m[5] + p[2] * m[6] + m[7]
    z = p[0] * m[8] + p[1] * m[9] + p[2] * m[10] + m[11]
    return np.array([x, y, z], F)          # affine light transforms: w == 1


def np_transform_vector(m, v):
    return np.array([v[0] * m[0] + v[1] * m[1] + v[2] * m[2],
                     v[0] * m[4] + v[1] * m[5] + v[2] * m[6],
                     v[0] * m[8] + v[1] * m[9] + v[2] * m[10]], F)


def np_quad(ray, q, limit):
    """Quad::GetIntersect (shape/quad.cpp:66-94). Returns t or None."""
    inv = np.asarray(q["inv"], F).reshape(16)
    with np.errstate(all="ignore"):
        o = np_transform_point(inv, ray[0:3].astype(F))
        d = np_transform_vector(inv, ray[4:7].astype(F))
        if d[1] == 0:
            return None
        t = F(F(-o[1]) / d[1])
        if t > limit or t <= ray[3] or t > ray[7]:
            return None
        px = F(o[0] + F(t * d[0]))
        pz = F(o[2] + F(t * d[2]))
    hx, hy = F(F(q["size_x"]) * F(0.5)), F(F(q["size_y"]) * F(0.5))
    if px > hx or px < -hx or pz > hy or pz < -hy:
        return None
    return t


def brute(pos, idx, quads, rays):
    """Closest (t, prim, u, v, number of primitives at that t) and any-hit by brute force."""
    tri = np.asarray(pos, F).reshape(-1, 3)[np.asarray(idx).reshape(-1, 3)]
    ntri = len(tri)
    out = []
    for r in rays:
        ok, t, u, v = np_triangles(r, tri)
        best = (FLT_MAX, 0xFFFFFFFF, F(0), F(0))
        any_hit = bool(ok.any())
        ts = []
        if any_hit:
            tt = np.where(ok, t, FLT_MAX)
            k = int(np.argmin(tt))               # lowest id among equal t
            best = (tt[k], k, u[k], v[k])
            ts = list(tt[ok])
        for qi, q in enumerate(quads or []):
            tq = np_quad(r, q, best[0])
            if tq is not None:
                any_hit = True
                ts.append(tq)
                best = (tq, ntri + qi, F(0), F(0))
        ties = sum(1 for x in ts if x == best[0]) if best[1] != 0xFFFFFFFF else 0
        out.append((best, ties, any_hit))
    return out


# ---- ray sets ---------------------------------------------------------------------------------

def mixed_rays(n, seed, lo=-0.3, hi=1.3):
    """Random origins and directions with a mix of windows: open (0, inf), bounce (1e-4, inf),
    shadow-like finite segments and narrow windows."""
    g = np.random.default_rng(seed)
    o = g.uniform(lo, hi, (n, 3)).astype(F)
    d = g.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(F)
    kind = g.integers(0, 4, n)
    tmin = np.where(kind == 1, 1e-4, np.where(kind == 3, g.uniform(0.1, 0.6, n), 0.0)).astype(F)
    tmax = np.where(kind == 2, g.uniform(0.05, 1.5, n),
                    np.where(kind == 3, tmin + g.uniform(0.01, 0.4, n), FLT_MAX)).astype(F)
    return np.concatenate([o, tmin[:, None], d, tmax[:, None]], 1).astype(F)


def axis_parallel_rays(n, seed):
    """Directions exactly along +-x, +-y, +-z (the 1e-5 clamp of resolveRayData and the parallel
    branch of the root slab) and directions with one or two zero components."""
    g = np.random.default_rng(seed)
    o = g.uniform(0.02, 0.98, (n, 3)).astype(F)
    d = np.zeros((n, 3), F)
    for i in range(n):
        k = i % 9
        if k < 6:
            d[i, k // 2] = 1.0 if k % 2 == 0 else -1.0
        else:
            a, b = [(0, 1), (1, 2), (0, 2)][k - 6]
            v = g.normal(size=2)
            v /= np.linalg.norm(v)
            d[i, a], d[i, b] = v
    tmin = np.zeros(n, F)
    tmax = np.full(n, FLT_MAX, F)
    return np.concatenate([o, tmin[:, None], d, tmax[:, None]], 1).astype(F)


def small_soup(n, seed):
    g = np.random.default_rng(seed)
    c = g.uniform(0, 1, (n, 3))
    s = np.exp(g.normal(np.log(0.08), 0.5, n))[:, None, None]
    a = g.normal(size=(n, 3, 3))
    a /= np.linalg.norm(a, axis=2, keepdims=True)
    a -= a.mean(axis=1, keepdims=True)
    tri = (c[:, None, :] + s * a).astype(F)
    return tri.reshape(-1, 3), np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)


def _check(pos, idx, quads, rays, width, allow_root_quirk=False):
    ob = OracleBvh(pos, idx, quads, width=width)
    tree = ob.closest(rays)
    tree_occ = ob.occluded(rays)
    cpp_brute = ob.brute_closest(rays)
    cpp_brute_occ = ob.brute_occluded(rays)
    ref = brute(pos, idx, quads, rays)
    n_hit = n_tie = 0
    for i, ((bt, bp, bu, bv), ties, anyh) in enumerate(ref):
        tt, tp = tree[i, 0], int(tree[i, 1].view(np.uint32))
        root_quirk = rays[i, 3] < 0
        if allow_root_quirk and root_quirk and tp == 0xFFFFFFFF:
            assert not tree_occ[i], i
            continue
        assert np.float32(tt).view(np.uint32) == np.float32(bt).view(np.uint32), (i, tt, bt, rays[i])
        if bp == 0xFFFFFFFF:
            assert tp == 0xFFFFFFFF, i
        else:
            n_hit += 1
            if ties > 1:
                n_tie += 1
            else:
                assert tp == bp, (i, tp, bp)
            if tp == bp:
                assert tree[i, 2].view(np.uint32) == F(bu).view(np.uint32), i
                assert tree[i, 3].view(np.uint32) == F(bv).view(np.uint32), i
        assert bool(tree_occ[i]) == anyh, (i, tree_occ[i], anyh, rays[i])
        # the oracle's own C++ brute force agrees with the independent one
        assert cpp_brute[i, 0].view(np.uint32) == F(bt).view(np.uint32), i
        assert bool(cpp_brute_occ[i]) == anyh, i
    return n_hit, n_tie


@pytest.mark.parametrize("width", [8, 4])
def test_soup_closest_and_any_hit_equal_brute_force(width):
    pos, idx = small_soup(800, 11)
    rays = mixed_rays(1500, 12)
    n_hit, _ = _check(pos, idx, None, rays, width)
    assert n_hit > 300          # the set exercises real hits, not only misses


@pytest.mark.parametrize("width", [8, 4])
def test_soup_axis_parallel_rays_equal_brute_force(width):
    pos, idx = small_soup(600, 21)
    rays = axis_parallel_rays(900, 22)
    n_hit, _ = _check(pos, idx, None, rays, width)
    assert n_hit > 100


@pytest.mark.parametrize("width", [8, 4])
def test_cornell_room_with_quad_equals_brute_force(width):
    sc = scenes.cornell_sphere()
    rays = np.concatenate([mixed_rays(700, 31, lo=0.01, hi=0.99), axis_parallel_rays(270, 32)])
    # rays aimed at the light quad from below so the quad is hit often
    g = np.random.default_rng(33)
    o = g.uniform(0.1, 0.9, (200, 3)).astype(F)
    o[:, 1] = g.uniform(0.05, 0.5, 200)
    tgt = np.stack([g.uniform(0.35, 0.65, 200), np.full(200, 0.999), g.uniform(0.35, 0.65, 200)], 1)
    d = (tgt - o)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(F)
    up = np.concatenate([o, np.zeros((200, 1)), d, np.full((200, 1), FLT_MAX)], 1).astype(F)
    rays = np.concatenate([rays, up]).astype(F)
    n_hit, _ = _check(sc["pos"], sc["idx"], sc["quads"], rays, width)
    ntri = len(sc["idx"])
    quad_hits = OracleBvh(sc["pos"], sc["idx"], sc["quads"], width=width).closest(up)[:, 1].view(np.uint32)
    assert (quad_hits >= ntri).sum() > 20 and (quad_hits != 0xFFFFFFFF).sum() > 150
    assert n_hit > 800


@pytest.mark.parametrize("width", [8, 4])
def test_shared_edges_give_exact_ties(width):
    """A grid of quads split into triangle pairs: rays through shared edges and vertices hit
    several triangles at one t (watertightness); the tree must return that t and one of them."""
    n = 8
    xs = np.linspace(0, 1, n + 1, dtype=F)
    tris = []
    for i in range(n):
        for j in range(n):
            a = (xs[i], xs[j], F(0.5))
            b = (xs[i + 1], xs[j], F(0.5))
            c = (xs[i + 1], xs[j + 1], F(0.5))
            e = (xs[i], xs[j + 1], F(0.5))
            tris += [a, b, c, a, c, e]
    pos = np.array(tris, F)
    idx = np.arange(len(pos), dtype=np.uint32).reshape(-1, 3)
    g = np.random.default_rng(41)
    k = g.integers(1, n, (300, 2))
    o = np.stack([xs[k[:, 0]], xs[k[:, 1]], np.full(300, -1.0, F)], 1).astype(F)
    o[150:, 0] = g.uniform(0.05, 0.95, 150)       # half through edges only
    d = np.tile(np.array([0, 0, 1], F), (300, 1))
    rays = np.concatenate([o, np.zeros((300, 1)), d, np.full((300, 1), FLT_MAX)], 1).astype(F)
    n_hit, n_tie = _check(pos, idx, None, rays, width)
    assert n_hit == 300 and n_tie > 150


@pytest.mark.parametrize("width", [8, 4])
def test_negative_tmin_root_quirk_and_non_finite_rays(width):
    pos, idx = small_soup(400, 51)
    rays = mixed_rays(400, 52, lo=0.2, hi=0.8)
    rays[:200, 3] = -0.5                          # window starts below zero, origin inside the box
    bad = mixed_rays(60, 53)
    bad[0:10, 0] = np.nan
    bad[10:20, 4] = np.nan
    bad[20:30, 5] = np.inf
    bad[30:40, 1] = -np.inf
    bad[40:50, 7] = np.nan
    bad[50:60, 3] = np.nan
    _check(pos, idx, None, np.concatenate([rays, bad]).astype(F), width, allow_root_quirk=True)
    # the quirk itself: every origin-inside ray with tmin < 0 is a miss for the tree walk
    ob = OracleBvh(pos, idx, None, width=width)
    got = ob.closest(rays[:200])
    assert (got[:, 1].view(np.uint32) == 0xFFFFFFFF).all()
    assert not ob.occluded(rays[:200]).any()


def test_plausible_mistakes_are_caught():
    """The brute force distinguishes the reference's det order and window from plausible slips:
    (e0 + e1) + e2 instead of e0 + (e1 + e2) changes some t bits; a closed lower bound accepts a
    hit exactly at tmin; a t recomputed with one FMA-free order swap differs somewhere."""
    pos, idx = small_soup(300, 61)
    rays = mixed_rays(600, 62)
    tri = np.asarray(pos, F).reshape(-1, 3)[idx]
    diffs = 0
    for r in rays:
        ok, t, _, _ = np_triangles(r, tri)
        kx, ky, kz, sx, sy, sz = np_prepare(r[4:7])
        with np.errstate(all="ignore"):
            p = (tri - r[None, None, 0:3]).astype(F)
            px, py, pz = p[:, :, kx], p[:, :, ky], p[:, :, kz]
            px = (py * sx + px).astype(F)
            pz = (py * sz + pz).astype(F)
            e0 = (px[:, 1] * pz[:, 2] - pz[:, 1] * px[:, 2]).astype(F)
            e1 = (px[:, 2] * pz[:, 0] - pz[:, 2] * px[:, 0]).astype(F)
            e2 = (px[:, 0] * pz[:, 1] - pz[:, 0] * px[:, 1]).astype(F)
            det_alt = ((e0 + e1) + e2).astype(F)
            py = (py * sy).astype(F)
            t_alt = (((e0 * py[:, 0] + e1 * py[:, 1]) + e2 * py[:, 2]) * (F(1) / det_alt)).astype(F)
        diffs += int((ok & (t_alt.view(np.uint32) != t.view(np.uint32))).sum())
    assert diffs > 0
    # an exact hit at tmin: the reference rejects it (t > tmin), a closed window would accept it
    tri1 = np.array([[[-1, -1, 2], [1, -1, 2], [0, 1, 2]]], F)
    r = np.array([0, 0, 0, 2, 0, 0, 1, FLT_MAX], F)
    ok, t, _, _ = np_triangles(r, tri1)
    assert t[0] == 2 and not ok[0]
    r[3] = np.nextafter(F(2), F(0))
    ok, _, _, _ = np_triangles(r, tri1)
    assert ok[0]
    ob = OracleBvh(tri1.reshape(-1, 3), np.array([[0, 1, 2]], np.uint32))
    assert ob.closest(np.array([[0, 0, 0, 2, 0, 0, 1, FLT_MAX]], F))[0, 1].view(np.uint32) == 0xFFFFFFFF
    assert ob.closest(r[None])[0, 0] == 2


def test_sse_unordered_neq_agrees_with_avx_ordered_neq():
    """The 4-wide restatement uses SSE's unordered det != 0 (simd_intel.h:137-138), the 8-wide one
    AVX's ordered _CMP_NEQ_OQ (simd_intel.h:271-272). They can only differ for a NaN det, which the
    preceding edge-sign mask already excludes (all >= 0 or all <= 0 rules out NaN, and a sum of
    same-signed non-NaN values is not NaN). Triangles and rays whose products overflow to +-inf or
    give NaN edge functions exercise exactly those lanes: both widths give the same bits."""
    g = np.random.default_rng(71)
    n = 200
    c = g.uniform(-1, 1, (n, 3)) * np.where(g.random((n, 1)) < 0.5, 1e19, 1.0)
    a = g.normal(size=(n, 3, 3)) * np.where(g.random((n, 1, 1)) < 0.3, 1e19, 0.3)
    assert np.isfinite(c).all() and np.isfinite(a).all()
    tri = (c[:, None, :] + a).astype(F)
    tri[::17, 0, 0] = F(1e20)                    # finite vertices (the build needs finite boxes)
    tri[::23, 1, 2] = F(-1e20)                   # whose edge products overflow to +-inf / NaN
    pos = tri.reshape(-1, 3)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    rays = mixed_rays(400, 72, lo=-2, hi=2)
    rays[::7, 0:3] *= F(1e19)
    rays[::11, 4:7] = np.array([1e-30, 1.0, 1e30], F)
    b8 = OracleBvh(pos, idx, None, width=8)
    b4 = OracleBvh(pos, idx, None, width=4)
    h8, h4 = b8.brute_closest(rays), b4.brute_closest(rays)
    assert np.array_equal(h8.view(np.uint32), h4.view(np.uint32))
    assert np.array_equal(b8.brute_occluded(rays), b4.brute_occluded(rays))
    assert (h8[:, 1].view(np.uint32) != 0xFFFFFFFF).sum() > 5
