"""GPU BSSRDF multi-hit query (sortgpu_trace_multi) against the oracle: the four records (slot
order included) and the counts bitwise equal, for both node widths, on dense soups where most
rays have more than four candidate hits (so the reference's record replacement and its lazily
resolved maxt decide the result), on the Cornell room (quads are never reported) and on the
edge cases."""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleBvh
from sort_amd import Obvh, Qbvh, SortGpuError, scenes

pytestmark = pytest.mark.gpu


def _dense_soup(n, seed, size=0.08):
    g = np.random.default_rng(seed)
    c = g.random((n, 1, 3)).astype(np.float32)
    a = g.normal(size=(n, 3, 3)).astype(np.float32)
    v = (c + np.float32(size) * a).astype(np.float32)
    return np.ascontiguousarray(v.reshape(-1, 3)), np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)


def _check(acc, ob, rays, ray_mat, tri_mat, label):
    print(f"[multi] {label} width={acc.width} rays={len(rays)}", flush=True)
    hits, cnt = acc.multi(rays, ray_mat, tri_mat)
    ref, ref_cnt = ob.multi(rays, ray_mat, tri_mat, threads=16)
    bad = np.nonzero(np.any(hits.reshape(len(rays), -1).view(np.uint32) != ref.reshape(len(rays), -1).view(np.uint32),
                            axis=1) | (cnt != ref_cnt))[0]
    assert bad.size == 0, f"{label}: {bad.size} mismatches, first {bad[:5]}: gpu {hits[bad[0]]} {cnt[bad[0]]} " \
                          f"ref {ref[bad[0]]} {ref_cnt[bad[0]]}"
    return hits, cnt


@pytest.mark.parametrize("width", [8, 4])
def test_soup_multi_parity(gpu_ctx, width):
    pos, idx = _dense_soup(1 << 15, 51, size=0.05)
    g = np.random.default_rng(52)
    tri_mat = g.integers(0, 3, len(idx)).astype(np.uint32)
    acc = (Obvh if width == 8 else Qbvh)(gpu_ctx).build(pos, idx)
    ob = OracleBvh(pos, idx, width=width)
    rays = scenes.random_rays(1 << 16, 53, lo=-0.2, hi=1.2)
    ray_mat = g.integers(0, 3, len(rays)).astype(np.uint32)
    _, cnt = _check(acc, ob, rays, ray_mat, tri_mat, "dense soup")
    assert np.mean(cnt == 4) > 0.3 and np.mean(cnt == 0) > 0.01
    # short probes (the BSSRDF's sampling segments) from surface points
    cam = scenes.pinhole_rays(256, 256)
    hits = acc.get_intersect(cam)
    probes = scenes.incoherent_rays(cam, hits, pos, idx, count=1 << 15)
    probes[:, 3] = np.float32(-0.05)
    probes[:, 7] = np.float32(0.1)
    _check(acc, ob, probes, g.integers(0, 3, len(probes)).astype(np.uint32), tri_mat, "probe segments")
    # one material everywhere
    _check(acc, ob, rays, np.zeros(len(rays), np.uint32), np.zeros(len(idx), np.uint32), "single material")


@pytest.mark.parametrize("width", [8, 4])
def test_cornell_multi_parity(gpu_ctx, width):
    sc = scenes.cornell_sphere()
    acc = (Obvh if width == 8 else Qbvh)(gpu_ctx).build(sc["pos"], sc["idx"], sc["quads"])
    ob = OracleBvh(sc["pos"], sc["idx"], sc["quads"], width=width)
    rays = scenes.random_rays(1 << 15, 61, lo=0.01, hi=0.99)
    ray_mat = np.random.default_rng(62).integers(0, 5, len(rays)).astype(np.uint32)
    hits, cnt = _check(acc, ob, rays, ray_mat, sc["tri_material"], "cornell")
    prims = hits[:, :, 1].copy().view(np.uint32)
    assert not np.any(prims[prims != 0xFFFFFFFF] >= sc["ntris"])  # never a quad


def test_multi_edge_cases(gpu_ctx):
    pos, idx = _dense_soup(4096, 71)
    tri_mat = np.zeros(len(idx), np.uint32)
    acc = Obvh(gpu_ctx).build(pos, idx)
    rays = scenes.random_rays(2048, 72, lo=-0.2, hi=1.2)
    with pytest.raises(SortGpuError, match="triangle materials"):
        acc.multi(rays, np.zeros(len(rays), np.uint32), tri_mat[:-1])
    h, c = acc.multi(np.zeros((0, 8), np.float32), np.zeros(0, np.uint32), tri_mat)
    assert h.shape == (0, 4, 4) and c.shape == (0,)
    empty = Obvh(gpu_ctx).build(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    h, c = empty.multi(rays, np.zeros(len(rays), np.uint32), np.zeros(1, np.uint32))
    assert np.all(c == 0) and np.all(h[:, :, 0] == np.float32(3.4028235e38))
    # a material no triangle has: nothing, and device inputs on a side stream
    s = torch.cuda.Stream()
    d_rays = torch.from_numpy(rays).cuda()
    d_rm = torch.full((len(rays),), 7, dtype=torch.int32, device="cuda")
    d_tm = torch.from_numpy(tri_mat.view(np.int32)).cuda()
    torch.cuda.synchronize()
    h, c = acc.multi(d_rays, d_rm, d_tm, stream=s.cuda_stream)
    s.synchronize()
    assert int(c.sum()) == 0


def test_multi_order_changes_no_result(gpu_ctx):
    """Batches of 2^18 probes or more are walked in origin-cell order by default; never (1), always
    (2) and the default (0) give the same bits, equal to the oracle's."""
    from sort_amd import batch
    pos, idx = _dense_soup(1 << 14, 81, size=0.05)
    tri_mat = np.random.default_rng(82).integers(0, 2, len(idx)).astype(np.uint32)
    acc = Obvh(gpu_ctx).build(pos, idx)
    ob = OracleBvh(pos, idx, width=8)
    rays = scenes.random_rays(300_000, 83, lo=-0.2, hi=1.2)
    ray_mat = np.random.default_rng(84).integers(0, 2, len(rays)).astype(np.uint32)
    res = {}
    try:
        for mode in (1, 2, 0):
            batch.set_tuning(gpu_ctx, reorder=mode)
            print(f"[multi order] mode {mode}", flush=True)
            res[mode] = acc.multi(rays, ray_mat, tri_mat)
    finally:
        batch.set_tuning(gpu_ctx)
    for mode in (2, 0):
        assert np.array_equal(res[mode][0].view(np.uint32), res[1][0].view(np.uint32))
        assert np.array_equal(res[mode][1], res[1][1])
    ref, ref_cnt = ob.multi(rays, ray_mat, tri_mat, threads=16)
    assert np.array_equal(res[0][0].view(np.uint32), ref.view(np.uint32)) and np.array_equal(res[0][1], ref_cnt)
