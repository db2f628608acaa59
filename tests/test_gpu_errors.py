"""Error flags of the traversal kernels: raised by the kernel, reported by the take calls
(sortgpu_accel_errors / sortgpu_pool_errors) with a message naming the flag, cleared when taken,
and never left behind to fail a later, clean call.

Flags are forced from the host with legal inputs only: an attenuation chain longer than the
walk's 4,096-surface cap, and a pooled ray whose scene index lies outside the pool (both are
checked cases of the kernels, which then write a defined result and exit normally).
"""
import numpy as np
import pytest
import torch

from sort_amd import Obvh, SortGpuError, scenes
from sort_amd import _lib
from sort_amd.batch import Pool

pytestmark = pytest.mark.gpu


def _stacked_planes(n):
    """n parallel squares |x|, |y| <= 1 at z = 0.01 k, two triangles each."""
    z = (np.arange(n, dtype=np.float32) * np.float32(0.01))
    corners = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], np.float32)
    pos = np.zeros((n, 4, 3), np.float32)
    pos[:, :, :2] = corners
    pos[:, :, 2] = z[:, None]
    base = (4 * np.arange(n, dtype=np.uint32))[:, None]
    idx = np.concatenate([base + np.array([0, 1, 2], np.uint32), base + np.array([0, 2, 3], np.uint32)])
    return np.ascontiguousarray(pos.reshape(-1, 3)), np.ascontiguousarray(idx)


def test_attenuation_cap_flag_reported_once_and_cleared(gpu_ctx):
    pos, idx = _stacked_planes(4200)
    acc = Obvh(gpu_ctx)
    acc.build(pos, idx)
    tr = np.ones((len(idx), 4), np.float32)  # fully clear: the colour never reaches black
    rays = np.array([[0.1, 0.2, -1.0, 0.0, 0.0, 0.0, 1.0, 1e3],     # crosses 4,200 surfaces
                     [0.1, 0.2, -1.0, 0.0, 0.0, 0.0, 1.0, 1.005]],  # crosses 1 surface
                    np.float32)
    print("[errors] attenuation cap: 8,400 triangles, 2 rays", flush=True)
    assert acc.check_errors() == 0
    att = acc.attenuation(rays, tr)
    assert att[0, 3] == 4096.0 and np.all(att[0, :3] == 1.0)
    assert att[1, 3] == 1.0
    with pytest.raises(SortGpuError, match="4096 surfaces") as ei:
        acc.check_errors()
    assert "stack overflow" not in str(ei.value)
    # taken: the next check and a later counters call on the same accelerator are clean
    assert acc.check_errors() == 0
    hits, st = acc.counters(rays)
    assert st["inner_nodes"] > 0
    assert hits[0, 0] == np.float32(1.0)  # the plane at z = 0


def test_pool_bad_scene_flag(gpu_ctx):
    sc = scenes.cornell_sphere()
    acc = Obvh(gpu_ctx)
    acc.build(sc["pos"], sc["idx"], sc["quads"])
    pool = Pool(gpu_ctx, [acc])
    rays = scenes.pinhole_rays(16, 16, eye=(0.5, 0.5, 0.02), target=(0.5, 0.5, 1.0), vfov_deg=70.0)
    d_rays = torch.from_numpy(rays).to("cuda:0")
    good = torch.zeros(len(rays), dtype=torch.int32, device="cuda:0")
    print("[errors] pool scene ids: 256 rays, one pool scene", flush=True)
    ref = pool.trace(d_rays, good).cpu().numpy()
    assert pool.check_errors() == 0
    bad = good.clone()
    bad[7] = 5  # outside the one-scene pool
    out = pool.trace(d_rays, bad).cpu().numpy()
    with pytest.raises(SortGpuError, match="outside the pool"):
        pool.check_errors()
    assert pool.check_errors() == 0
    assert out[7, 1].view(np.uint32) == 0xFFFFFFFF and out[7, 0] == np.finfo(np.float32).max
    keep = np.arange(len(rays)) != 7
    assert np.array_equal(out[keep].view(np.uint32), ref[keep].view(np.uint32))
    assert _lib.FLAG_BAD_SCENE == 2
