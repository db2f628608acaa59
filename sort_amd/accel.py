"""Python mirror of SORT's spatial-accelerator plugin surface over the C ABI.

``Obvh`` / ``Qbvh`` stand where SORT registers ``DEFINE_RTTI(Obvh, Accelerator)`` /
``DEFINE_RTTI(Qbvh, Accelerator)`` (accel/fast_bvh.h:103-108):

* ``Serialize`` parameters ``max_node_depth`` / ``max_pri_in_leaf`` (fast_bvh.h:159-162) are
  constructor arguments (defaults 28 / 16, the Blender exporter's values);
* ``build(...)`` replaces ``Accelerator::Build(const Scene&)`` (accelerator.h:128);
* ``get_intersect(rays)`` replaces ``Accelerator::GetIntersect`` (accelerator.h:74) for a batch;
* ``is_occluded(rays)`` replaces ``Accelerator::IsOccluded`` (accelerator.h:85), i.e. the live
  ``GetIntersect(query_shadow = true)`` path (fast_bvh.hpp:362-373).

Rays are (n, 8) float32 ``[ox, oy, oz, tmin, dx, dy, dz, tmax]``; hits are (n, 4) records
``[t, prim (uint32 bits), u, v]``. Device tensors (torch, on the context's device) are used in
place; numpy arrays are copied to the device and results copied back.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib

MISS = 0xFFFFFFFF


class Context:
    """Binds the library to one HIP device (``device < 0``: host-only, builds only)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().sortgpu_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            lib().sortgpu_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _quads_array(quads) -> tuple[ctypes.Array | None, int]:
    if quads is None or len(quads) == 0:
        return None, 0
    arr = (_lib.Quad * len(quads))()
    for i, q in enumerate(quads):
        mat = np.asarray(q["mat"], np.float32).reshape(16)
        inv = np.asarray(q["inv"], np.float32).reshape(16)
        arr[i].mat[:] = mat.tolist()
        arr[i].inv[:] = inv.tolist()
        arr[i].size_x = float(q["size_x"])
        arr[i].size_y = float(q["size_y"])
    return arr, len(quads)


class Accelerator:
    """Fast BVH over triangles + area-light quads, resident in HBM."""

    width = 8

    def __init__(self, ctx: Context, max_node_depth: int = 28, max_pri_in_leaf: int = 16,
                 build_threads: int = 0):
        self.ctx = ctx
        self.max_node_depth = max_node_depth
        self.max_pri_in_leaf = max_pri_in_leaf
        self.build_threads = build_threads
        self.handle = None
        self.ntris = 0
        self.nquads = 0

    # -- Build -------------------------------------------------------------------------------
    def build(self, positions, indices, quads=None):
        """positions (V,3) float32, indices (T,3) uint32, quads: list of dicts with mat, inv,
        size_x, size_y. Primitive ids: triangles 0..T-1 then quads T..T+Q-1."""
        self.release()
        pos = np.ascontiguousarray(np.asarray(positions, np.float32).reshape(-1, 3))
        idx = np.ascontiguousarray(np.asarray(indices, np.uint32).reshape(-1, 3))
        qa, nq = _quads_array(quads)
        opts = _lib.BuildOpts(self.width, self.max_node_depth, self.max_pri_in_leaf, self.build_threads)
        h = ctypes.c_void_p()
        check(lib().sortgpu_accel_build(
            self.ctx.handle, pos.ctypes.data if len(pos) else None, len(pos),
            idx.ctypes.data if len(idx) else None, len(idx),
            ctypes.cast(qa, ctypes.c_void_p) if qa is not None else None, nq,
            ctypes.byref(opts), ctypes.byref(h)), self.ctx.handle)
        self.handle = h
        self.ntris = len(idx)
        self.nquads = nq
        return self

    def release(self):
        if self.handle:
            lib().sortgpu_accel_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    # -- inspection ----------------------------------------------------------------------------
    @property
    def valid(self) -> bool:
        return bool(lib().sortgpu_accel_valid(self.handle))

    def bbox(self) -> np.ndarray:
        out = np.zeros(6, np.float32)
        check(lib().sortgpu_accel_bbox(self.handle, out.ctypes.data))
        return out

    def info(self) -> dict:
        out = np.zeros(8, np.uint64)
        check(lib().sortgpu_accel_info(self.handle, out.ctypes.data))
        keys = ["width", "depth", "inner_nodes", "leaves", "prims", "max_stack", "device_bytes", "build_ms"]
        return {k: int(v) for k, v in zip(keys, out)}

    def prim_order(self) -> np.ndarray:
        n = self.info()["prims"]
        out = np.zeros(n, np.uint32)
        check(lib().sortgpu_accel_prim_order(self.handle, out.ctypes.data if n else None, n))
        return out

    def tree_dump(self) -> np.ndarray:
        n = check(lib().sortgpu_accel_tree_dump(self.handle, None, 0))
        out = np.zeros(n, np.uint32)
        check(lib().sortgpu_accel_tree_dump(self.handle, out.ctypes.data if n else None, n))
        return out

    # -- queries -------------------------------------------------------------------------------
    def _device_rays(self, rays):
        import torch
        if isinstance(rays, torch.Tensor):
            assert rays.is_cuda and rays.dtype == torch.float32 and rays.is_contiguous()
            return rays.view(-1, 8), False
        arr = np.ascontiguousarray(np.asarray(rays, np.float32).reshape(-1, 8))
        return torch.from_numpy(arr).to(f"cuda:{self.ctx.device}"), True

    def get_intersect(self, rays, stream=None, out=None):
        """Closest hits, (n,4) float32 records [t, prim bits, u, v] (written into `out`, a
        contiguous (n,4) float32 device tensor, when given)."""
        import torch
        d_rays, host = self._device_rays(rays)
        n = d_rays.shape[0]
        if out is not None:
            assert out.is_cuda and out.dtype == torch.float32 and out.is_contiguous() and out.numel() == 4 * n
        hits = out if out is not None else torch.empty((n, 4), dtype=torch.float32, device=d_rays.device)
        s = stream if stream is not None else torch.cuda.current_stream(d_rays.device).cuda_stream
        check(lib().sortgpu_trace_closest(self.handle, d_rays.data_ptr(), n, hits.data_ptr(), s),
              self.ctx.handle)
        return hits.cpu().numpy() if host else hits

    def is_occluded(self, rays, stream=None, out=None):
        """Any-hit: uint8 per ray, 1 when some primitive lies in (tmin, tmax] (written into `out`,
        a contiguous (n,) uint8 device tensor, when given)."""
        import torch
        d_rays, host = self._device_rays(rays)
        n = d_rays.shape[0]
        if out is not None:
            assert out.is_cuda and out.dtype == torch.uint8 and out.is_contiguous() and out.numel() == n
        occ = out if out is not None else torch.empty((n,), dtype=torch.uint8, device=d_rays.device)
        s = stream if stream is not None else torch.cuda.current_stream(d_rays.device).cuda_stream
        check(lib().sortgpu_trace_any(self.handle, d_rays.data_ptr(), n, occ.data_ptr(), s), self.ctx.handle)
        return occ.cpu().numpy() if host else occ

    def attenuation(self, rays, transparency, stream=None):
        """Transparent-shadow attenuation (SORT's Scene::GetAttenuation): (n,4) float32 records
        [r, g, b, surfaces passed]. `transparency`: (n_prims, 4) float32, per primitive id
        (triangles then quads) the material's transparency r, g, b and 1.0 when the material has
        transparency (else 0.0)."""
        import torch
        d_rays, host = self._device_rays(rays)
        n = d_rays.shape[0]
        if isinstance(transparency, torch.Tensor):
            tr = transparency
            assert tr.is_cuda and tr.dtype == torch.float32 and tr.is_contiguous()
        else:
            tr = torch.from_numpy(np.ascontiguousarray(np.asarray(transparency, np.float32).reshape(-1, 4))).to(
                d_rays.device)
        att = torch.empty((n, 4), dtype=torch.float32, device=d_rays.device)
        s = stream if stream is not None else torch.cuda.current_stream(d_rays.device).cuda_stream
        check(lib().sortgpu_trace_attenuation(self.handle, d_rays.data_ptr(), n, tr.data_ptr(), tr.numel() // 4,
                                              att.data_ptr(), s), self.ctx.handle)
        return att.cpu().numpy() if host else att

    def multi(self, rays, ray_material, tri_material, stream=None):
        """BSSRDF multi-hit query (SORT's GetIntersect with BSSRDFIntersections): ((n,4,4) float32
        records [t, prim bits, u, v] in slot order, (n,) uint32 counts). ``ray_material``: (n,)
        material per ray; ``tri_material``: material per triangle id."""
        import torch
        d_rays, host = self._device_rays(rays)
        n = d_rays.shape[0]
        dev = d_rays.device

        def u32(x):
            if isinstance(x, torch.Tensor):
                assert x.is_cuda and x.dtype == torch.int32 and x.is_contiguous()
                return x
            return torch.from_numpy(np.ascontiguousarray(np.asarray(x, np.uint32).view(np.int32))).to(dev)

        rm, tm = u32(ray_material), u32(tri_material)
        hits = torch.empty((n, 4, 4), dtype=torch.float32, device=dev)
        cnt = torch.empty((n,), dtype=torch.int32, device=dev)
        s = stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream
        check(lib().sortgpu_trace_multi(self.handle, d_rays.data_ptr(), n, rm.data_ptr(), tm.data_ptr(), tm.numel(),
                                        hits.data_ptr(), cnt.data_ptr(), s), self.ctx.handle)
        if host:
            return hits.cpu().numpy(), cnt.cpu().numpy().view(np.uint32)
        return hits, cnt

    def counters(self, rays, mode: int = 0):
        """(results, {inner_nodes, leaves, tris, quads, wide_nodes}) for closest (0) or any-hit (1);
        wide_nodes = tests of 8-wide nodes with more than four children."""
        import torch
        d_rays, host = self._device_rays(rays)
        n = d_rays.shape[0]
        out = np.zeros(5, np.uint64)
        s = torch.cuda.current_stream(d_rays.device).cuda_stream
        if mode == 0:
            res = torch.empty((n, 4), dtype=torch.float32, device=d_rays.device)
            check(lib().sortgpu_trace_counters(self.handle, d_rays.data_ptr(), n, 0, res.data_ptr(), None,
                                               out.ctypes.data, s), self.ctx.handle)
        else:
            res = torch.empty((n,), dtype=torch.uint8, device=d_rays.device)
            check(lib().sortgpu_trace_counters(self.handle, d_rays.data_ptr(), n, 1, None, res.data_ptr(),
                                               out.ctypes.data, s), self.ctx.handle)
        st = dict(zip(["inner_nodes", "leaves", "tris", "quads", "wide_nodes"], [int(x) for x in out]))
        return (res.cpu().numpy() if host else res), st


    def check_errors(self, stream=None) -> int:
        """Takes the error flags the accelerator's kernels raised since the last check (after the
        work queued on `stream`); raises SortGpuError naming each flag when any was raised, else
        returns 0."""
        import torch
        flags = ctypes.c_uint32(0)
        s = stream if stream is not None else torch.cuda.current_stream(self.ctx.device).cuda_stream
        check(lib().sortgpu_accel_errors(self.handle, s, ctypes.byref(flags)), self.ctx.handle)
        return flags.value


class Obvh(Accelerator):
    """8-wide SAH BVH (SORT's Obvh, accel/obvh.h)."""
    width = 8


class Qbvh(Accelerator):
    """4-wide SAH BVH (SORT's Qbvh, accel/qbvh.h)."""
    width = 4


def hit_prims(hits) -> np.ndarray:
    """Primitive ids (uint32) of (n,4) hit records."""
    h = np.ascontiguousarray(np.asarray(hits if isinstance(hits, np.ndarray) else hits.cpu().numpy(), np.float32))
    return h[:, 1].view(np.uint32)
