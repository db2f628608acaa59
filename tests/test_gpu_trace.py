"""GPU parity of the traversal kernels (closest hit + any hit) against the oracle.

The bar is stricter than the nominal one (prim ids equal except near-ties, t within 1e-5
relative, any-hit bitwise): because the HIP kernels reproduce the reference's arithmetic and
child order exactly, every field of every hit record must be bitwise equal to the oracle's.
"""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleBvh
from sort_amd import MISS, Obvh, Qbvh, scenes

pytestmark = pytest.mark.gpu

FLT_MAX = np.float32(3.4028234663852886e38)


def _pair(ctx, pos, idx, quads=None, width=8, **kw):
    cls = Obvh if width == 8 else Qbvh
    acc = cls(ctx, **{k: v for k, v in kw.items() if k in ("max_node_depth", "max_pri_in_leaf")})
    acc.build(pos, idx, quads)
    ob = OracleBvh(pos, idx, quads, width=width, max_depth=kw.get("max_node_depth", 28),
                   max_leaf=kw.get("max_pri_in_leaf", 16))
    return acc, ob


def _check(acc, ob, rays, label="", window_property=True):
    print(f"[parity] {label} width={acc.width} rays={len(rays)}", flush=True)
    d_rays = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    hits = acc.get_intersect(d_rays).cpu().numpy()
    occ = acc.is_occluded(d_rays).cpu().numpy()
    ref = ob.closest(rays, threads=16)
    ref_occ = ob.occluded(rays, threads=16)
    bad = np.nonzero(np.any(hits.view(np.uint32) != ref.view(np.uint32), axis=1))[0]
    assert bad.size == 0, f"{label}: {bad.size} closest mismatches, first {bad[:5]}: gpu {hits[bad[:3]]} ref {ref[bad[:3]]}"
    bad = np.nonzero(occ != ref_occ)[0]
    assert bad.size == 0, f"{label}: {bad.size} any-hit mismatches, first {bad[:5]}"
    if window_property:
        # size-independent property: occluded <=> a closest hit exists in the same window
        assert np.array_equal(occ.astype(bool), hits[:, 1].view(np.uint32) != MISS)
    return hits, occ


@pytest.mark.parametrize("width", [8, 4])
def test_cornell_scene_parity(gpu_ctx, width):
    sc = scenes.cornell_sphere()
    acc, ob = _pair(gpu_ctx, sc["pos"], sc["idx"], sc["quads"], width)
    cam = scenes.pinhole_rays(128, 128, eye=(0.5, 0.5, 0.02), target=(0.5, 0.5, 1.0), vfov_deg=70.0)
    cam_hits, _ = _check(acc, ob, cam, "cornell camera")
    assert np.mean(cam_hits[:, 1].view(np.uint32) != MISS) > 0.99  # closed room
    rnd = scenes.random_rays(20000, 101, lo=0.01, hi=0.99)
    hits, _ = _check(acc, ob, rnd, "cornell random")
    assert np.any(hits[:, 1].view(np.uint32) == sc["ntris"])  # some rays hit the light quad
    # bounce rays from the camera hits
    inc = scenes.incoherent_rays(cam, cam_hits, sc["pos"], sc["idx"], count=20000)
    _check(acc, ob, inc, "cornell bounce")


@pytest.mark.parametrize("width", [8, 4])
def test_soup_parity_all_ray_kinds(gpu_ctx, width):
    pos, idx = scenes.soup(1 << 16, seed=11)
    acc, ob = _pair(gpu_ctx, pos, idx, width=width)
    cam = scenes.pinhole_rays(256, 256)
    hits, _ = _check(acc, ob, cam, "soup coherent")
    inc = scenes.incoherent_rays(cam, hits, pos, idx, count=1 << 16)
    _check(acc, ob, inc, "soup incoherent")
    sh = scenes.shadow_rays(cam, hits, pos, idx, count=1 << 16)
    _check(acc, ob, sh, "soup shadow")
    _check(acc, ob, scenes.random_rays(1 << 15, 5), "soup random")


def test_counters_match_oracle_statistics(gpu_ctx):
    pos, idx = scenes.soup(1 << 15, seed=12)
    acc, ob = _pair(gpu_ctx, pos, idx)
    rays = scenes.random_rays(1 << 14, 9, lo=0.0, hi=1.0)
    d = torch.from_numpy(rays).cuda()
    _, st = acc.counters(d, mode=0)
    ref = {}
    ob.closest(rays, stats=ref)
    assert (st["inner_nodes"], st["leaves"], st["tris"], st["quads"]) == \
        (ref["inner_nodes"], ref["leaves"], ref["tris"], ref["quads"])
    _, st = acc.counters(d, mode=1)
    ref = {}
    ob.occluded(rays, stats=ref)
    # any-hit leaves after the first blocker: nodes and leaves visited match exactly
    assert (st["inner_nodes"], st["leaves"]) == (ref["inner_nodes"], ref["leaves"])


def test_edge_cases(gpu_ctx):
    rays = scenes.random_rays(4096, 21, lo=-0.2, hi=1.2)
    # empty scene: valid handle, everything misses
    acc = Obvh(gpu_ctx).build(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    assert not acc.valid
    h = acc.get_intersect(rays)
    assert np.all(h[:, 1].view(np.uint32) == MISS) and np.all(h[:, 0] == FLT_MAX)
    assert not np.any(acc.is_occluded(rays))
    # zero rays
    assert acc.get_intersect(np.zeros((0, 8), np.float32)).shape == (0, 4)
    # single triangle: the root is a leaf
    pos = np.array([[0, 0, 0.5], [1, 0, 0.5], [0, 1, 0.5]], np.float32)
    acc, ob = _pair(gpu_ctx, pos, np.array([[0, 1, 2]], np.uint32))
    _check(acc, ob, rays, "single triangle")
    # 200 triangles sharing one centroid: the partition degenerates into one big leaf
    g = np.random.default_rng(4)
    a = g.normal(size=(200, 3, 3)).astype(np.float32) * 0.1
    a -= a.mean(axis=1, keepdims=True)
    a += np.float32(0.5)
    pos, idx = scenes.tris_to_mesh(list(a))
    acc, ob = _pair(gpu_ctx, pos, idx)
    _check(acc, ob, scenes.random_rays(8192, 22, lo=0.3, hi=0.7), "shared centroid")
    # quad-only scene and a forced shallow tree (max depth 2 -> big leaves)
    sc = scenes.cornell_sphere()
    acc, ob = _pair(gpu_ctx, np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32), sc["quads"])
    _check(acc, ob, scenes.random_rays(4096, 23, lo=0.3, hi=0.7), "quad only")
    acc, ob = _pair(gpu_ctx, sc["pos"], sc["idx"], sc["quads"], max_node_depth=2)
    _check(acc, ob, scenes.random_rays(8192, 24, lo=0.01, hi=0.99), "depth 2")
    acc, ob = _pair(gpu_ctx, sc["pos"], sc["idx"], sc["quads"], max_pri_in_leaf=1)
    _check(acc, ob, scenes.random_rays(8192, 25, lo=0.01, hi=0.99), "leaf 1")


def test_axis_parallel_and_windowed_rays(gpu_ctx):
    pos, idx = scenes.soup(1 << 14, seed=13)
    acc, ob = _pair(gpu_ctx, pos, idx)
    g = np.random.default_rng(6)
    n = 6 * 2048
    o = g.uniform(-0.1, 1.1, (n, 3)).astype(np.float32)
    d = np.zeros((n, 3), np.float32)
    for k in range(6):
        d[k * 2048:(k + 1) * 2048, k % 3] = 1.0 if k < 3 else -1.0
    d[::7] += np.float32(3e-6)  # below the 1e-5 clamp on the other axes
    neg0 = np.zeros(n, bool)
    neg0[::11] = True
    neg0[np.arange(n) // 2048 % 3 == 0] = False  # keep x as the major axis where it is one
    d[neg0, 0] = np.float32(-0.0)                # sign(-0) = +1 in the clamp
    tmin = g.uniform(0, 0.3, n).astype(np.float32)
    tmax = (tmin + g.uniform(0, 1.0, n)).astype(np.float32)
    _check(acc, ob, scenes.make_rays(o, d, tmin, tmax), "axis parallel windowed")


@pytest.mark.parametrize("width", [8, 4])
def test_select_form_slab_tests(gpu_ctx, width):
    """Rays with infinite or NaN components, and every ray of a tree whose box reaches past 1e12,
    take the x86-select form of the slab test (the IEEE min/max form is only used where no slab
    distance can be NaN); both forms must give the oracle's bits. A ray's results must not depend
    on which form its wave neighbours use, so ordinary and non-finite rays are interleaved."""
    pos, idx = scenes.soup(1 << 13, seed=31)
    acc, ob = _pair(gpu_ctx, pos, idx, width=width)
    rays = scenes.random_rays(16384, 32, lo=-0.2, hi=1.2)
    inf, nan = np.float32(np.inf), np.float32(np.nan)
    r = rays.copy()
    n = len(r)
    k = np.arange(n)
    r[k % 8 == 1, 0] = inf                       # infinite origin
    r[k % 8 == 2, 5] = -inf                      # infinite direction
    r[k % 8 == 3, 1] = nan                       # NaN origin
    r[k % 8 == 4, 6] = nan                       # NaN direction
    sel = k % 8 == 5                             # infinite origin and direction on one axis
    r[sel, 2] = -inf
    r[sel, 6] = inf
    r[k % 8 == 6, 7] = nan                       # NaN tmax
    r[k % 16 == 7, 3] = nan                      # NaN tmin
    _check(acc, ob, r, "non-finite rays", window_property=False)
    # the same soup and rays scaled by 2e12: a normal tree (SAH areas and the triangle test's
    # t numerator stay finite) whose box is past the IEEE bound, so every ray takes the select form
    s = np.float32(2e12)
    acc2, ob2 = _pair(gpu_ctx, (pos * s).astype(np.float32), idx, width=width)
    assert acc2.info()["inner_nodes"] > 100
    far = rays.copy()
    far[:, :3] *= s
    far[:, 7] = np.float32(3e38)
    hits, _ = _check(acc2, ob2, far, "scaled by 2e12")
    assert np.mean(hits[:, 1].view(np.uint32) != MISS) > 0.01


def test_traversal_order_changes_no_result(gpu_ctx):
    """Large batches are traced in coherence order (sorted by origin and direction cell); the
    order must not change a single bit of any result: never (1), always (2) and the default (0,
    which sorts these 300K-ray batches) agree with each other and with the oracle, also for a
    batch whose rays all share one sort key and for two streams tracing one accelerator at once
    (each with its own stream-ordered sort buffers)."""
    from sort_amd import batch
    pos, idx = scenes.soup(1 << 16, seed=41)
    acc, ob = _pair(gpu_ctx, pos, idx)
    rays = scenes.random_rays(300_000, 42, lo=-0.2, hi=1.2)
    one_key = np.repeat(rays[7:8], 1 << 18, axis=0)
    d_rays, d_one = torch.from_numpy(rays).cuda(), torch.from_numpy(one_key).cuda()
    res = {}
    try:
        for mode in (1, 2, 0):
            batch.set_tuning(gpu_ctx, reorder=mode)
            print(f"[order] mode {mode}", flush=True)
            res[mode] = [acc.get_intersect(d_rays).cpu().numpy(), acc.is_occluded(d_rays).cpu().numpy(),
                         acc.get_intersect(d_one).cpu().numpy(), acc.is_occluded(d_one).cpu().numpy()]
        print("[order] two streams", flush=True)
        s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
        h1 = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
        h2 = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
        back = torch.from_numpy(np.ascontiguousarray(rays[::-1])).cuda()
        torch.cuda.synchronize()
        for _ in range(3):
            acc.get_intersect(d_rays, out=h1, stream=s1.cuda_stream)
            acc.get_intersect(back, out=h2, stream=s2.cuda_stream)
        torch.cuda.synchronize()
        two = [h1.cpu().numpy(), h2.cpu().numpy()[::-1]]
    finally:
        batch.set_tuning(gpu_ctx)
    for mode in (2, 0):
        for a, b in zip(res[mode], res[1]):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), mode
    for h in two:
        assert np.array_equal(h.view(np.uint32), res[1][0].view(np.uint32))
    ref = ob.closest(rays, threads=16)
    assert np.array_equal(res[0][0].view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(res[0][1], ob.occluded(rays, threads=16))


@pytest.mark.parametrize("width", [8, 4])
def test_full_size_soup_parity(gpu_ctx, width):
    """The 1M-triangle workload at full size: 4M coherent, 4M incoherent and 4M shadow rays."""
    pos, idx = scenes.soup()
    acc, ob = _pair(gpu_ctx, pos, idx, width=width)
    cam = scenes.pinhole_rays(2048, 2048)
    hits, _ = _check(acc, ob, cam, "C2 coherent")
    inc = scenes.incoherent_rays(cam, hits, pos, idx)
    _check(acc, ob, inc, "C2 incoherent")
    sh = scenes.shadow_rays(cam, hits, pos, idx)
    _check(acc, ob, sh, "C2 shadow")
