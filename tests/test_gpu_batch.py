"""Batched form: one persistent launch over a concatenated pool of many scenes must give, for
every ray, exactly the hit its own scene's accelerator gives (which is bitwise the oracle's),
for both node widths and every tuning of the kernel."""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleBvh
from sort_amd import MISS, Context, batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def work8(gpu_ctx):
    print("[batch] building scaled workload (16 scenes)", flush=True)
    return batch.build_workload(gpu_ctx, n_scenes=16, width=8, scale=0.05)


def _per_scene(w, d_rays, any_hit):
    outs = []
    for i, acc in enumerate(w["accels"]):
        lo, hi = w["offsets"][i], w["offsets"][i + 1]
        r = d_rays[lo:hi]
        outs.append(acc.is_occluded(r) if any_hit else acc.get_intersect(r))
    return torch.cat(outs).cpu().numpy()


@pytest.mark.parametrize("any_hit", [False, True])
def test_pool_equals_per_scene_and_oracle(gpu_ctx, work8, any_hit):
    w = work8
    pool = batch.Pool(gpu_ctx, w["accels"])
    d_rays = torch.from_numpy(w["rays"]).cuda()
    d_sid = torch.from_numpy(w["scene_of_ray"].view(np.int32)).cuda()
    got = pool.trace(d_rays, d_sid, any_hit=any_hit).cpu().numpy()
    ref = _per_scene(w, d_rays, any_hit)
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
    # oracle on three scenes' rays
    for i in (0, 7, 15):
        lo, hi = w["offsets"][i], w["offsets"][i + 1]
        pos, idx = w["geometry"][i]
        ob = OracleBvh(pos, idx)
        o = ob.occluded(w["rays"][lo:hi], threads=16) if any_hit else ob.closest(w["rays"][lo:hi], threads=16)
        assert np.array_equal(got[lo:hi].view(np.uint8), o.view(np.uint8)), i


def test_pool_shuffled_scene_order_and_bad_ids(gpu_ctx, work8):
    w = work8
    pool = batch.Pool(gpu_ctx, w["accels"])
    g = np.random.default_rng(1)
    perm = g.permutation(len(w["rays"]))
    rays = np.ascontiguousarray(w["rays"][perm])
    sid = np.ascontiguousarray(w["scene_of_ray"][perm])
    sid_bad = sid.copy()
    sid_bad[::97] = 1000  # out-of-range scene ids miss
    d_rays = torch.from_numpy(rays).cuda()
    got = pool.trace(d_rays, torch.from_numpy(sid.view(np.int32)).cuda()).cpu().numpy()
    ref = _per_scene(w, torch.from_numpy(w["rays"]).cuda(), False)[perm]
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    bad = pool.trace(d_rays, torch.from_numpy(sid_bad.view(np.int32)).cuda()).cpu().numpy()
    assert np.all(bad[::97, 1].view(np.uint32) == MISS)
    keep = np.ones(len(sid), bool)
    keep[::97] = False
    assert np.array_equal(bad[keep].view(np.uint32), ref[keep].view(np.uint32))


def test_tunings_and_widths_give_identical_hits(gpu_ctx, work8):
    w = work8
    d_rays = torch.from_numpy(w["rays"]).cuda()
    d_sid = torch.from_numpy(w["scene_of_ray"].view(np.int32)).cuda()
    pool = batch.Pool(gpu_ctx, w["accels"])
    ref = pool.trace(d_rays, d_sid).cpu().numpy()
    ref_any = pool.trace(d_rays, d_sid, any_hit=True).cpu().numpy()
    try:
        for refill, lds, bpc, lb, lm in [(8, 8, 0, 0, 0), (32, 16, 2, 64, 1), (64, 8, 1, 1, 1), (1, 16, 0, 16, 8),
                                         (16, 16, 0, 48, 32), (16, 8, 0, 64, 64)]:
            batch.set_tuning(gpu_ctx, refill, lds, bpc, lb, lm)
            got = pool.trace(d_rays, d_sid).cpu().numpy()
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (refill, lds, bpc, lb, lm)
            occ = pool.trace(d_rays, d_sid, any_hit=True).cpu().numpy()
            assert np.array_equal(occ, ref_any), (refill, lds, bpc, lb, lm)
    finally:
        batch.set_tuning(gpu_ctx, 0, 0, 0)
    # 4-wide trees over the same rays: the per-triangle arithmetic does not depend on the tree,
    # so hit t agrees bitwise except where a box test of one tree culls a grazing hit the other
    # keeps (float slab tests are not exactly conservative) -- allowed for 1e-5 of the rays
    w4 = batch.build_workload(gpu_ctx, n_scenes=16, width=4, scale=0.05)
    pool4 = batch.Pool(gpu_ctx, w4["accels"])
    got4 = pool4.trace(d_rays, d_sid).cpu().numpy()
    same = got4[:, 0].view(np.uint32) == ref[:, 0].view(np.uint32)
    assert same.mean() >= 1 - 1e-5, (~same).sum()


@pytest.mark.parametrize("any_hit", [False, True])
def test_pool_traversal_order_changes_no_result(gpu_ctx, work8, any_hit):
    """A pooled batch taken in (scene, origin-cell) order gives every ray the same result as in
    index order, including rays whose scene id is outside the pool (they sort last and miss)."""
    w = work8
    pool = batch.Pool(gpu_ctx, w["accels"])
    d_rays = torch.from_numpy(w["rays"]).cuda()
    sid = w["scene_of_ray"].copy()
    sid[::101] = 1 << 20  # out of range
    d_sid = torch.from_numpy(sid.view(np.int32)).cuda()
    print(f"[batch] order never / always over {len(sid)} rays, any={any_hit}", flush=True)
    try:
        batch.set_tuning(gpu_ctx, reorder=1)
        plain = pool.trace(d_rays, d_sid, any_hit=any_hit).cpu().numpy()
        batch.set_tuning(gpu_ctx, reorder=2)
        ordered = pool.trace(d_rays, d_sid, any_hit=any_hit).cpu().numpy()
    finally:
        batch.set_tuning(gpu_ctx)
    assert np.array_equal(plain.view(np.uint8), ordered.view(np.uint8))
    with pytest.raises(Exception, match="outside the pool"):
        pool.check_errors()
    pool.close()


def test_pool_of_many_scenes_orders_on_a_coarser_grid(gpu_ctx):
    """A pool of more scenes than the default key leaves room for (2,100 ids need 12 bits, so the
    origin cell drops to a 64^3 grid) is still taken in (scene, cell) order, and every ray's hit
    equals index order's and its own scene's accelerator's."""
    from sort_amd import Obvh, scenes
    n_scenes = 2100
    g = np.random.default_rng(2100)
    accels = []
    print(f"[batch] building {n_scenes} tiny scenes", flush=True)
    for i in range(n_scenes):
        c = g.random((6, 1, 3)).astype(np.float32)
        pos = (c + np.float32(0.3) * g.normal(size=(6, 3, 3)).astype(np.float32)).reshape(-1, 3)
        accels.append(Obvh(gpu_ctx).build(np.ascontiguousarray(pos), np.arange(18, dtype=np.uint32).reshape(6, 3)))
    pool = batch.Pool(gpu_ctx, accels)
    rays = scenes.random_rays(1 << 14, 77, lo=0.0, hi=1.0)
    sid = g.integers(0, n_scenes, len(rays)).astype(np.uint32)
    d_rays = torch.from_numpy(rays).cuda()
    d_sid = torch.from_numpy(sid.view(np.int32)).cuda()
    try:
        batch.set_tuning(gpu_ctx, reorder=1)
        plain = pool.trace(d_rays, d_sid).cpu().numpy()
        batch.set_tuning(gpu_ctx, reorder=2)
        ordered = pool.trace(d_rays, d_sid).cpu().numpy()
    finally:
        batch.set_tuning(gpu_ctx)
    assert np.array_equal(plain.view(np.uint32), ordered.view(np.uint32))
    for i in range(0, n_scenes, 97):
        m = sid == i
        if m.any():
            own = accels[i].get_intersect(d_rays[torch.from_numpy(np.nonzero(m)[0]).cuda()]).cpu().numpy()
            assert np.array_equal(own.view(np.uint32), plain[m].view(np.uint32))
    assert np.any(plain[:, 1].view(np.uint32) != MISS)
    pool.close()
    for a in accels:
        a.release()
