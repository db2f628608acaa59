"""Batched form: many scenes' rays traced by one persistent launch over a concatenated pool of
their trees (``sortgpu_pool_*``), and the 64-scene workload it is measured on.

Workload (seeded): 64 triangle soups generated like the 1M-triangle soup, with triangle counts
~ LogNormal(ln 150K, 1.0) clipped to [10K, 2M] (median 150K); per scene an incoherent ray batch
~ LogNormal(ln 256K, 1.0) clipped to [16K, 4M] rays, made from that scene's own camera hits.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib, scenes
from ._lib import check, lib
from .accel import Context, Obvh, Qbvh


class Pool:
    """Concatenated device pool of built accelerators (one node width)."""

    def __init__(self, ctx: Context, accels):
        self.ctx = ctx
        self.accels = list(accels)
        arr = (ctypes.c_void_p * max(1, len(self.accels)))(*[a.handle.value for a in self.accels])
        h = ctypes.c_void_p()
        check(lib().sortgpu_pool_create(ctx.handle, ctypes.cast(arr, ctypes.c_void_p), len(self.accels),
                                        ctypes.byref(h)), ctx.handle)
        self.handle = h

    def info(self):
        out = np.zeros(4, np.uint64)
        check(lib().sortgpu_pool_info(self.handle, out.ctypes.data))
        return dict(zip(["scenes", "width", "device_bytes", "max_stack"], [int(x) for x in out]))

    def trace(self, d_rays, d_scene, any_hit=False, out=None, stream=None):
        import torch
        n = d_rays.shape[0]
        if out is None:
            out = (torch.empty((n,), dtype=torch.uint8, device=d_rays.device) if any_hit
                   else torch.empty((n, 4), dtype=torch.float32, device=d_rays.device))
        s = stream if stream is not None else torch.cuda.current_stream(d_rays.device).cuda_stream
        check(lib().sortgpu_pool_trace(self.handle, d_rays.data_ptr(), d_scene.data_ptr(), n, 1 if any_hit else 0,
                                       out.data_ptr(), s), self.ctx.handle)
        return out

    def check_errors(self, stream=None) -> int:
        """Takes the error flags the pool's traces raised since the last check; raises
        SortGpuError naming each flag when any was raised (e.g. a ray naming a scene outside the
        pool), else returns 0."""
        import torch
        flags = ctypes.c_uint32(0)
        s = stream if stream is not None else torch.cuda.current_stream(self.ctx.device).cuda_stream
        check(lib().sortgpu_pool_errors(self.handle, s, ctypes.byref(flags)), self.ctx.handle)
        return flags.value

    def close(self):
        if self.handle:
            lib().sortgpu_pool_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_tuning(ctx: Context, refill=0, lds_entries=0, blocks_per_cu=0, leaf_batch=0, leaf_min=0, reorder=0,
               shade_order=0):
    """Traversal knobs of the context (0 = default); they change speed only, never a hit.
    reorder: 0 = sort large trace batches into coherence order (default), 1 = never, 2 = always.
    shade_order: 0 = shade each bounce's hits of waves of 2^18+ paths in spatial order (default),
    1 = never, 2 = always."""
    t = _lib.Tuning(refill, lds_entries, blocks_per_cu, leaf_batch, leaf_min, reorder, shade_order)
    check(lib().sortgpu_ctx_set_tuning(ctx.handle, ctypes.byref(t)), ctx.handle)


def workload_sizes(n_scenes=64, seed=scenes.SEED_BATCH):
    g = scenes.rng(seed)
    tris = np.clip(np.exp(g.normal(np.log(150_000), 1.0, n_scenes)), 10_000, 2_000_000).astype(np.int64)
    rays = np.clip(np.exp(g.normal(np.log(262_144), 1.0, n_scenes)), 16_384, 4_194_304).astype(np.int64)
    return tris, rays


def build_workload(ctx: Context, n_scenes=64, width=8, seed=scenes.SEED_BATCH, scale=1.0):
    """Builds the scenes (on the device context `ctx`) and their incoherent rays. Returns
    dict(accels, geometry, rays (N,8) float32, scene_of_ray (N,) uint32, offsets, tris, nrays).
    `scale` shrinks the triangle and ray counts (tests)."""
    import torch
    tris, nrays = workload_sizes(n_scenes, seed)
    tris = np.maximum((tris * scale).astype(np.int64), 64)
    nrays = np.maximum((nrays * scale).astype(np.int64), 256)
    cls = Obvh if width == 8 else Qbvh
    accels, geo, ray_chunks, sid = [], [], [], []
    for i in range(n_scenes):
        pos, idx = scenes.soup(int(tris[i]), seed=seed * 1000 + i)
        acc = cls(ctx).build(pos, idx)
        cam = scenes.pinhole_rays(256, 256)  # camera hits seed the incoherent origins
        hits = acc.get_intersect(torch.from_numpy(cam).to(f"cuda:{ctx.device}")).cpu().numpy()
        r = scenes.incoherent_rays(cam, hits, pos, idx, count=int(nrays[i]), seed=seed * 1000 + 500 + i)
        accels.append(acc)
        geo.append((pos, idx))
        ray_chunks.append(r)
        sid.append(np.full(len(r), i, np.uint32))
    rays = np.concatenate(ray_chunks)
    scene_of_ray = np.concatenate(sid)
    offsets = np.concatenate([[0], np.cumsum([len(c) for c in ray_chunks])]).astype(np.int64)
    return {"accels": accels, "geometry": geo, "rays": rays, "scene_of_ray": scene_of_ray, "offsets": offsets,
            "tris": tris, "nrays": nrays}
