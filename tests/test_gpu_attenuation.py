"""GPU parity of the transparent-shadow attenuation walk (Scene::GetAttenuation,
core/scene.cpp:192-208, over Accelerator::GetAttenuation, accel/accelerator.cpp:26-68) against the
oracle: colour and surface count bitwise equal on every ray, for both node widths, on triangle and
quad scenes, with the table-size refusal of the C-ABI.
"""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleBvh
from sort_amd import Obvh, Qbvh, SortGpuError, scenes

pytestmark = pytest.mark.gpu


def _pair(ctx, pos, idx, quads=None, width=8):
    acc = (Obvh if width == 8 else Qbvh)(ctx)
    acc.build(pos, idx, quads)
    return acc, OracleBvh(pos, idx, quads, width=width)


def _table(n, seed, frac_transparent=0.8, lo=0.2):
    g = np.random.default_rng(seed)
    tr = np.zeros((n, 4), np.float32)
    tr[:, :3] = g.uniform(lo, 1.0, (n, 3)).astype(np.float32)
    tr[:, 3] = (g.random(n) < frac_transparent).astype(np.float32)
    return tr


def _check(acc, ob, rays, tr, label):
    print(f"[atten] {label} width={acc.width} rays={len(rays)}", flush=True)
    got = acc.attenuation(rays, tr)
    ref = ob.attenuation(rays, tr, threads=16)
    bad = np.nonzero(np.any(got.view(np.uint32) != ref.view(np.uint32), axis=1))[0]
    assert bad.size == 0, f"{label}: {bad.size} mismatches, first {bad[:5]}: gpu {got[bad[:3]]} ref {ref[bad[:3]]}"
    return got


def _dense_soup(n, seed, size=0.08):
    g = np.random.default_rng(seed)
    c = g.random((n, 1, 3)).astype(np.float32)
    a = g.normal(size=(n, 3, 3)).astype(np.float32)
    v = (c + np.float32(size) * a).astype(np.float32)
    return np.ascontiguousarray(v.reshape(-1, 3)), np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)


@pytest.mark.parametrize("width", [8, 4])
def test_soup_attenuation_parity(gpu_ctx, width):
    pos, idx = _dense_soup(1 << 14, 71)
    acc, ob = _pair(gpu_ctx, pos, idx, width=width)
    rays = scenes.random_rays(1 << 16, 72, lo=-0.2, hi=1.2)
    got = _check(acc, ob, rays, _table(len(idx), 73), "dense soup mixed")
    assert np.mean(got[:, 3] >= 4) > 0.1 and np.mean(np.all(got[:, :3] == 0, axis=1)) > 0.1
    # every material transparent: long chains, nothing blocked by an opaque surface
    got = _check(acc, ob, rays, _table(len(idx), 74, frac_transparent=1.0, lo=0.9), "dense soup clear")
    assert got[:, 3].max() >= 20
    # every material opaque: black exactly where the any-hit query reports an occluder
    opaque = _table(len(idx), 75, frac_transparent=0.0)
    got = _check(acc, ob, rays, opaque, "dense soup opaque")
    occ = acc.is_occluded(rays).astype(bool)
    assert np.all(got[occ, :3] == 0) and np.all(got[~occ, :3] == 1) and np.all(got[:, 3] == 0)
    # the headline soup generator at 256K triangles, area-light shadow segments: few surfaces each
    pos, idx = scenes.soup(1 << 18, seed=76)
    acc, ob = _pair(gpu_ctx, pos, idx, width=width)
    cam = scenes.pinhole_rays(256, 256)
    hits = acc.get_intersect(cam)
    sh = scenes.shadow_rays(cam, hits, pos, idx, count=1 << 16)
    _check(acc, ob, sh, _table(len(idx), 77, frac_transparent=0.5), "soup shadow")


@pytest.mark.parametrize("width", [8, 4])
def test_quad_scene_attenuation_parity(gpu_ctx, width):
    """Cornell room (triangles + quads): quads are their own groups; the light quad is transparent."""
    sc = scenes.cornell_sphere()
    acc, ob = _pair(gpu_ctx, sc["pos"], sc["idx"], sc["quads"], width)
    n = sc["ntris"] + len(sc["quads"])
    rays = scenes.random_rays(1 << 15, 81, lo=0.01, hi=0.99)
    tr = _table(n, 82, frac_transparent=0.7)
    _check(acc, ob, rays, tr, "cornell mixed")
    tr[:] = [0.5, 0.75, 0.25, 1.0]
    got = _check(acc, ob, rays, tr, "cornell all transparent")
    assert got[:, 3].max() >= 2
    # a table longer than the primitive count is accepted; the extra records are never read
    _check(acc, ob, rays, np.concatenate([tr, np.full((16, 4), np.nan, np.float32)]), "cornell long table")


def test_attenuation_edge_cases(gpu_ctx):
    pos = np.array([[0, 0, 0.5], [1, 0, 0.5], [0, 1, 0.5]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    acc, ob = _pair(gpu_ctx, pos, idx)
    rays = scenes.random_rays(4096, 91, lo=-0.2, hi=1.2)
    _check(acc, ob, rays, np.array([[0.25, 0.5, 1.0, 1.0]], np.float32), "single triangle")
    # zero rays
    assert acc.attenuation(np.zeros((0, 8), np.float32), np.ones((1, 4), np.float32)).shape == (0, 4)
    # a table shorter than the primitive count is refused before any launch
    sc = scenes.cornell_sphere()
    acc2, _ = _pair(gpu_ctx, sc["pos"], sc["idx"], sc["quads"])
    with pytest.raises(SortGpuError, match="transparency records"):
        acc2.attenuation(rays, np.ones((sc["ntris"], 4), np.float32))  # quads missing
    # empty scene: nothing on any segment
    acc3 = Obvh(gpu_ctx).build(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    got = acc3.attenuation(rays, np.ones((1, 4), np.float32))
    assert np.all(got == np.array([1, 1, 1, 0], np.float32))
    # device inputs on a side stream
    s = torch.cuda.Stream()
    d_rays = torch.from_numpy(rays).cuda()
    d_tr = torch.tensor([[0.25, 0.5, 1.0, 1.0]], dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    got = acc.attenuation(d_rays, d_tr, stream=s.cuda_stream)
    s.synchronize()
    assert np.array_equal(got.cpu().numpy().view(np.uint32),
                          ob.attenuation(rays, d_tr.cpu().numpy()).view(np.uint32))


def test_attenuation_order_changes_no_result(gpu_ctx):
    """Batches of 2^18 rays or more are walked in origin-cell order by default; never (1), always
    (2) and the default (0) give the same bits, equal to the oracle's."""
    from sort_amd import batch
    pos, idx = _dense_soup(1 << 14, 95)
    acc, ob = _pair(gpu_ctx, pos, idx)
    rays = scenes.random_rays(300_000, 96, lo=-0.2, hi=1.2)
    tr = _table(len(idx), 97)
    d_rays, d_tr = torch.from_numpy(rays).cuda(), torch.from_numpy(tr).cuda()
    res = {}
    try:
        for mode in (1, 2, 0):
            batch.set_tuning(gpu_ctx, reorder=mode)
            print(f"[atten order] mode {mode}", flush=True)
            res[mode] = acc.attenuation(d_rays, d_tr).cpu().numpy()
    finally:
        batch.set_tuning(gpu_ctx)
    for mode in (2, 0):
        assert np.array_equal(res[mode].view(np.uint32), res[1].view(np.uint32)), mode
    assert np.array_equal(res[0].view(np.uint32), ob.attenuation(rays, tr, threads=16).view(np.uint32))
