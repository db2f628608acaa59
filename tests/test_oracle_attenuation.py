"""The oracle's transparent-shadow attenuation (Scene::GetAttenuation, core/scene.cpp:192-208, around
Accelerator::GetAttenuation, accel/accelerator.cpp:26-68, with the transparent-shadow walk of
fast_bvh.hpp:352-405), pinned without retyping it:

* closed forms: rays through stacked transparent planes get the product of the planes'
  transparencies, in order, and pass exactly that many surfaces; an opaque plane anywhere on the
  segment gives black; a black or colourless (no-transparency) material gives black;
* an all-opaque table reduces to the any-hit query (already pinned by brute force): black exactly
  where the ray is occluded, else white;
* an independent loop built only from the brute-force closest hit (every primitive tested, no
  tree) gives the same colour bits and surface counts on a soup mixing opaque and transparent
  triangles, wherever the outcome does not hang on the tree's visiting order.
"""
import numpy as np

from oracle.oracle import OracleBvh
from sort_amd import scenes


def _planes(zs, size=2.0):
    """Axis-aligned squares z = zs[k] over [-size/2+0.5, size/2+0.5]^2, two triangles each."""
    tris = []
    h = size / 2
    for z in zs:
        a, b, c, d = [0.5 - h, 0.5 - h, z], [0.5 + h, 0.5 - h, z], [0.5 + h, 0.5 + h, z], [0.5 - h, 0.5 + h, z]
        tris += [[a, b, c], [a, c, d]]
    return scenes.tris_to_mesh([np.array(t, np.float32) for t in tris])


def _down_rays(n, seed, tmax=np.float32(3e38)):
    g = np.random.default_rng(seed)
    o = np.column_stack([g.uniform(0.1, 0.9, n), g.uniform(0.1, 0.9, n), np.full(n, -1.0)]).astype(np.float32)
    d = np.column_stack([g.uniform(-0.05, 0.05, n), g.uniform(-0.05, 0.05, n), np.ones(n)]).astype(np.float32)
    return scenes.make_rays(o, d, np.float32(0.0), tmax)


def test_stacked_planes_closed_form():
    zs = [0.1, 0.3, 0.5, 0.7, 0.9]
    pos, idx = _planes(zs)
    cols = np.array([[0.5, 0.25, 1.0], [0.9, 0.8, 0.7], [0.6, 1.0, 0.5], [1.0, 0.3, 0.9], [0.75, 0.5, 0.25]],
                    np.float32)
    tr = np.zeros((len(idx), 4), np.float32)
    for k in range(len(zs)):
        tr[2 * k: 2 * k + 2, :3] = cols[k]
        tr[2 * k: 2 * k + 2, 3] = 1.0
    rays = _down_rays(512, 1)
    for width in (8, 4):
        ob = OracleBvh(pos, idx, width=width)
        out = ob.attenuation(rays, tr)
        want = np.ones(3, np.float32)
        for k in range(len(zs)):  # the planes are met in z order; products in that order
            want = want * cols[k]
        assert np.array_equal(out[:, :3], np.broadcast_to(want, (512, 3))), width
        assert np.all(out[:, 3] == len(zs))
        # a segment that ends between planes 3 and 4 (t < 1.6 from z = -1, direction ~ +z)
        short = rays.copy()
        short[:, 7] = np.float32(1.6)
        out = ob.attenuation(short, tr)
        want3 = cols[0] * cols[1] * cols[2]
        assert np.array_equal(out[:, :3], np.broadcast_to(want3, (512, 3)))
        assert np.all(out[:, 3] == 3)
        # an opaque third plane blocks; a black transparency blocks; a colourless material blocks
        for kind in ("opaque", "black", "colourless"):
            t2 = tr.copy()
            if kind == "opaque":
                t2[4:6] = [0.9, 0.9, 0.9, 0.0]
            elif kind == "black":
                t2[4:6] = [0.0, 0.0, 0.0, 1.0]
            else:
                t2[4:6, 3] = 0.0
            out = ob.attenuation(rays, t2)
            assert np.all(out[:, :3] == 0.0), kind
            assert np.all(out[:, 3] == 2), kind  # two transparent planes passed before it


def test_all_opaque_is_any_hit():
    pos, idx = scenes.soup(1 << 12, seed=51)
    rays = scenes.random_rays(4096, 52, lo=-0.2, hi=1.2)
    tr = np.zeros((len(idx), 4), np.float32)
    tr[:, :3] = 0.5  # colour ignored: no material has transparency
    for width in (8, 4):
        ob = OracleBvh(pos, idx, width=width)
        occ = ob.occluded(rays).astype(bool)
        out = ob.attenuation(rays, tr)
        assert np.all(out[occ, :3] == 0.0) and np.all(out[~occ, :3] == 1.0)
        assert np.all(out[:, 3] == 0)


def _brute_attenuation(pos, idx, rays, tr):
    """Scene::GetAttenuation with each step's hit from brute force over every primitive (no tree).

    Returns (colour, surfaces passed, ambiguous). A ray is ambiguous when some segment of its chain
    holds an opaque surface: the walk stops at an opaque surface as soon as it is the nearest hit
    found so far in its group, which depends on the order the tree visits them, so such a ray may
    end black where the chain, stepping past surfaces within 0.001 of a restart, does not."""
    full = OracleBvh(pos, idx, width=8)
    opq = np.nonzero(tr[:, 3] == 0.0)[0]
    opaque_only = OracleBvh(pos, np.ascontiguousarray(idx[opq]), width=8) if opq.size else None
    rays = rays.copy()
    att = np.ones((len(rays), 3), np.float32)
    steps = np.zeros(len(rays), np.int64)
    amb = np.zeros(len(rays), bool)
    done = np.zeros(len(rays), bool)
    for _ in range(64):
        act = np.nonzero(~done)[0]
        if act.size == 0:
            break
        if opaque_only is not None:
            amb[act] |= opaque_only.brute_closest(rays[act])[:, 1].view(np.uint32) != 0xFFFFFFFF
        h = full.brute_closest(rays[act])
        prim = h[:, 1].view(np.uint32)
        miss = prim == 0xFFFFFFFF
        done[act[miss]] = True
        act, h, prim = act[~miss], h[~miss], prim[~miss]
        has = tr[prim, 3] != 0.0
        a = np.where(has[:, None], tr[prim, :3], np.float32(0.0)).astype(np.float32)
        black = np.all(a == 0.0, axis=1)
        att[act[black]] = a[black]
        done[act[black]] = True
        act, h, a = act[~black], h[~black], a[~black]
        att[act] = att[act] * a
        steps[act] += 1
        done[act[np.all(att[act] == 0.0, axis=1)]] = True
        t = h[:, 0]
        o = rays[act, 0:3]
        d = rays[act, 4:7]
        rays[act, 0:3] = o + t[:, None] * d  # ray(t) = o + t * d, float32, no fused multiply-add
        rays[act, 3] = np.float32(0.001)
        rays[act, 7] = rays[act, 7] - t
    assert done.all()
    return att, steps, amb


def _dense_soup(n, seed, size=0.08):
    """Triangles of about `size` across with centres uniform in the unit cube: a ray through the
    cube passes several of them."""
    g = np.random.default_rng(seed)
    c = g.random((n, 1, 3)).astype(np.float32)
    a = g.normal(size=(n, 3, 3)).astype(np.float32)
    v = (c + np.float32(size) * a).astype(np.float32)
    return np.ascontiguousarray(v.reshape(-1, 3)), np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)


def test_soup_matches_brute_force_loop():
    pos, idx = _dense_soup(1 << 11, 61)
    g = np.random.default_rng(62)
    tr = np.zeros((len(idx), 4), np.float32)
    tr[:, :3] = g.uniform(0.2, 1.0, (len(idx), 3)).astype(np.float32)
    tr[:, 3] = (g.random(len(idx)) < 0.8).astype(np.float32)  # 80% of the triangles are transparent
    rays = scenes.random_rays(3000, 63, lo=-0.2, hi=1.2)
    for width in (8, 4):
        ob = OracleBvh(pos, idx, width=width)
        out = ob.attenuation(rays, tr)
        ref, steps, amb = _brute_attenuation(pos, idx, rays, tr)
        sure = ~amb
        assert np.array_equal(out[sure, :3].view(np.uint32), ref[sure].view(np.uint32)), width
        assert np.array_equal(out[sure, 3], steps[sure].astype(np.float32)), width
        # an ambiguous ray ends black, or as the chain does
        agree = np.all(out[amb, :3].view(np.uint32) == ref[amb].view(np.uint32), axis=1)
        assert np.all(agree | np.all(out[amb, :3] == 0.0, axis=1)), width
        assert (steps[sure] > 3).sum() > 100  # unambiguous rays passing several surfaces
        assert (np.all(out[:, :3] == 0, axis=1)).sum() > 100  # and blocked rays
