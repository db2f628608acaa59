"""Leaf records hold two triangles (A then B; an odd count ends on a one-triangle record), then
one quad per record. Every query must equal the oracle bit for bit whatever the leaves look like:
one-triangle leaves, odd and even counts, leaves of more than sixteen primitives (depth-limited
builds put every remaining primitive in one leaf), triangles followed by quads and quad-only
leaves — at both node widths. The visit statistics of closest hit must equal the oracle's too:
they count every triangle tested, so a record whose second slot is empty must not count it."""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleBvh
from sort_amd import MISS, Obvh, Qbvh, scenes

pytestmark = pytest.mark.gpu

# (max_node_depth, max_pri_in_leaf)
SHAPES = [(28, 1), (28, 2), (28, 3), (28, 5), (3, 16), (2, 16)]


def _scene():
    """The Cornell room with its sphere and light, plus eight small quads in a row under the
    ceiling, so that leaves hold quads next to triangles and quads alone."""
    sc = scenes.cornell_sphere()
    quads = list(sc["quads"])
    for k in range(8):
        quads.append(scenes.translate_quad((0.1 + 0.1 * k, 0.9, 0.3 + 0.05 * (k % 3)), 0.06, facing_down=True))
    return sc["pos"], sc["idx"], quads


def _rays():
    cam = scenes.pinhole_rays(64, 64, eye=(0.5, 0.5, 0.02), target=(0.5, 0.5, 1.0), vfov_deg=70.0)
    rnd = scenes.random_rays(6000, 7, lo=0.01, hi=0.99)
    up = scenes.random_rays(2000, 8, lo=0.05, hi=0.95)
    up[:, 4:7] = np.abs(up[:, 4:7]) * np.float32([0.1, 1.0, 0.1])  # towards the ceiling quads
    up[:, 4:7] /= np.linalg.norm(up[:, 4:7], axis=1, keepdims=True)
    return np.ascontiguousarray(np.concatenate([cam, rnd, up]).astype(np.float32))


@pytest.mark.parametrize("width", [8, 4])
@pytest.mark.parametrize("depth,leaf", SHAPES)
def test_leaf_shapes_parity(gpu_ctx, width, depth, leaf):
    pos, idx, quads = _scene()
    acc = (Obvh if width == 8 else Qbvh)(gpu_ctx, max_node_depth=depth, max_pri_in_leaf=leaf)
    acc.build(pos, idx, quads)
    ob = OracleBvh(pos, idx, quads, width=width, max_depth=depth, max_leaf=leaf)
    rays = _rays()
    print(f"[leaf] width={width} depth={depth} leaf={leaf} rays={len(rays)}", flush=True)
    d = torch.from_numpy(rays).cuda()
    hits = acc.get_intersect(d).cpu().numpy()
    ref = ob.closest(rays, threads=16)
    bad = np.nonzero(np.any(hits.view(np.uint32) != ref.view(np.uint32), axis=1))[0]
    assert bad.size == 0, f"closest: {bad.size} mismatches, first {bad[:5]}"
    prim = hits[:, 1].view(np.uint32)
    assert np.any((prim >= len(idx)) & (prim != MISS))  # some rays hit quads
    occ = acc.is_occluded(d).cpu().numpy()
    assert np.array_equal(occ.astype(bool), ob.occluded(rays, threads=16).astype(bool))
    # closest-hit visit statistics: nodes, leaves, triangles and quads tested
    _, st = acc.counters(d, mode=0)
    rst = {}
    ob.closest(rays, stats=rst)
    assert (st["inner_nodes"], st["leaves"], st["tris"], st["quads"]) == \
        (rst["inner_nodes"], rst["leaves"], rst["tris"], rst["quads"])
    # transparent-shadow walk: groups of W triangles end inside and across records
    g = np.random.default_rng(depth * 100 + leaf)
    n = len(idx) + len(quads)
    tr = np.zeros((n, 4), np.float32)
    tr[:, :3] = g.uniform(0.2, 1.0, (n, 3)).astype(np.float32)
    tr[:, 3] = (g.random(n) < 0.7).astype(np.float32)
    got = acc.attenuation(rays, tr)
    want = ob.attenuation(rays, tr, threads=16)
    bad = np.nonzero(np.any(got.view(np.uint32) != want.view(np.uint32), axis=1))[0]
    assert bad.size == 0, f"attenuation: {bad.size} mismatches, first {bad[:5]}"
    # multi-hit probes: four materials over the triangles
    tri_mat = g.integers(0, 4, len(idx)).astype(np.uint32)
    ray_mat = g.integers(0, 4, len(rays)).astype(np.uint32)
    mh, mc = acc.multi(rays, ray_mat, tri_mat)
    rh, rc = ob.multi(rays, ray_mat, tri_mat, threads=16)
    assert np.array_equal(mc, rc)
    assert np.array_equal(mh.reshape(len(rays), -1).view(np.uint32), rh.reshape(len(rays), -1).view(np.uint32))
