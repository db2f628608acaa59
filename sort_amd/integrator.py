"""Python mirror of SORT's path-tracing integrator entry points over the C ABI.

``PathTracing`` stands where SORT registers its ``PathTracing`` integrator
(integrator/pathtracing.h): ``max_recursive_depth`` (integrator.h:127) is a constructor argument,
``li(rays, keys)`` is the batched ``Integrator::Li`` (integrator.h:56) and ``render(...)`` runs the
tile loop of ``ImageEvaluation`` (work/image_evaluation/image_evaluation.cpp:184-264) as a
wavefront on the GPU.

Materials are plain parameter records (SORT builds the same BXDFs from TSL closures):
``lambert(rgb)``, ``ggx(rgb, roughness, eta, k)``, ``dielectric(reflectance, transmittance,
roughness, ior)``, ``disney(base, ...)``, ``transparent(attenuation)``. A scene with a transparent
material answers its shadow rays with SORT's transparent-shadow walk (Scene::GetAttenuation).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from ._lib import check, lib

LAMBERT, GGX, DIELECTRIC, DISNEY, TRANSPARENT = 0, 1, 2, 3, 4


def _mat(type_, base=(1, 1, 1), trans=(0, 0, 0), roughness=0.5, ior=1.5, eta=(1, 1, 1), k=(0, 0, 0), metallic=0.0,
         specular=0.5, specular_tint=0.0, sheen=0.0, sheen_tint=0.0, clearcoat=0.0, clearcoat_gloss=1.0):
    return dict(type=type_, base=base, trans=trans, roughness=roughness, ior=ior, eta=eta, k=k, metallic=metallic,
                specular=specular, specular_tint=specular_tint, sheen=sheen, sheen_tint=sheen_tint,
                clearcoat=clearcoat, clearcoat_gloss=clearcoat_gloss)


def lambert(rgb):
    return _mat(LAMBERT, base=rgb)


def ggx(rgb, roughness, eta=(0.27, 0.68, 1.22), k=(3.6, 2.6, 2.3)):
    """GGX microfacet reflection with a conductor Fresnel (defaults: copper-like eta / k)."""
    return _mat(GGX, base=rgb, roughness=roughness, eta=eta, k=k)


def dielectric(reflectance=(1, 1, 1), transmittance=(1, 1, 1), roughness=0.1, ior=1.5):
    return _mat(DIELECTRIC, base=reflectance, trans=transmittance, roughness=roughness, ior=ior)


def disney(base, metallic=0.0, specular=0.5, specular_tint=0.0, roughness=0.5, sheen=0.0, sheen_tint=0.5,
           clearcoat=0.0, clearcoat_gloss=1.0):
    return _mat(DISNEY, base=base, metallic=metallic, specular=specular, specular_tint=specular_tint,
                roughness=roughness, sheen=sheen, sheen_tint=sheen_tint, clearcoat=clearcoat,
                clearcoat_gloss=clearcoat_gloss)


def transparent(attenuation):
    """SORT's Transparent BXDF (the Transparent shader node, bsdf/transparent.cpp): light passes
    straight through, scaled by `attenuation`, both along paths and along shadow rays."""
    return _mat(TRANSPARENT, base=attenuation)


def material_array(mats):
    arr = (_lib.Material * max(1, len(mats)))()
    for i, m in enumerate(mats):
        r = arr[i]
        r.type = m["type"]
        r.base[:] = [float(x) for x in m["base"]]
        r.trans[:] = [float(x) for x in m["trans"]]
        r.roughness = m["roughness"]
        r.ior = m["ior"]
        r.eta[:] = [float(x) for x in m["eta"]]
        r.k[:] = [float(x) for x in m["k"]]
        for f in ("metallic", "specular", "specular_tint", "sheen", "sheen_tint", "clearcoat", "clearcoat_gloss"):
            setattr(r, f, float(m[f]))
    return arr


def camera(eye, target, up=(0, 1, 0), fov_y_deg=40.0):
    c = _lib.Camera()
    c.eye[:] = [float(x) for x in eye]
    c.target[:] = [float(x) for x in target]
    c.up[:] = [float(x) for x in up]
    c.fov_y = math.radians(fov_y_deg)
    return c


def scene_desc(materials, tri_material=None, normals=None, tangents=None, quad_radiance=None,
               env_radiance=(0.0, 0.0, 0.0), clamp=0.0):
    """A ``sortgpu_scene_desc`` and the arrays it points into (keep them alive while it is used).
    ``clamp`` > 0 clamps each sample's radiance to [0, clamp] (a scene file's ``clamping``)."""
    keep = []
    d = _lib.SceneDesc()

    def arr(a, dt):
        if a is None:
            return None
        x = np.ascontiguousarray(np.asarray(a, dt))
        keep.append(x)
        return x.ctypes.data

    d.normals = arr(normals, np.float32)
    d.tangents = arr(tangents, np.float32)
    d.tri_material = arr(tri_material, np.uint32)
    mats = material_array(materials)
    keep.append(mats)
    d.materials = ctypes.cast(mats, ctypes.c_void_p)
    d.nmaterials = len(materials)
    d.quad_radiance = arr(quad_radiance, np.float32)
    d.env_radiance[:] = [float(x) for x in env_radiance]
    d.clamp = float(clamp)
    return d, keep


class PathTracing:
    """Shading data + path tracer over a built accelerator (``sort_amd.accel.Accelerator``)."""

    def __init__(self, accel, materials, tri_material=None, normals=None, tangents=None, quad_radiance=None,
                 env_radiance=(0.0, 0.0, 0.0), max_recursive_depth=6, clamp=0.0):
        self.accel = accel
        self.max_recursive_depth = max_recursive_depth
        d, self._keep = scene_desc(materials, tri_material, normals, tangents, quad_radiance, env_radiance, clamp)
        h = ctypes.c_void_p()
        check(lib().sortgpu_scene_create(accel.handle, ctypes.byref(d), ctypes.byref(h)), accel.ctx.handle)
        self.handle = h

    def close(self):
        if self.handle:
            lib().sortgpu_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, cam, width, height, spp, seed=1, max_depth=None, out=None, stream=None):
        """Renders into a (height, width, 3) float32 device tensor (per-pixel mean radiance)."""
        import torch
        dev = f"cuda:{self.accel.ctx.device}"
        if out is None:
            out = torch.empty((height, width, 3), dtype=torch.float32, device=dev)
        s = stream if stream is not None else torch.cuda.current_stream(out.device).cuda_stream
        depth = self.max_recursive_depth if max_depth is None else max_depth
        check(lib().sortgpu_render(self.handle, ctypes.byref(cam), width, height, spp, depth, seed, out.data_ptr(), s),
              self.accel.ctx.handle)
        return out

    def ray_dump(self, cam, width, height, spp, pix_begin, pix_end, seed=1, max_depth=None, stream=None):
        """The extension rays (bounce 0 = camera ray) of the paths of pixels [pix_begin, pix_end)
        with their closest hits, slot ((pixel - pix_begin) * spp + sample) * max_depth + bounce.
        Returns device tensors rays (n, 8), hits (n, 4), valid (n,) uint8."""
        import torch
        dev = f"cuda:{self.accel.ctx.device}"
        depth = self.max_recursive_depth if max_depth is None else max_depth
        n = (pix_end - pix_begin) * spp * depth
        rays = torch.empty((n, 8), dtype=torch.float32, device=dev)
        hits = torch.empty((n, 4), dtype=torch.float32, device=dev)
        valid = torch.empty((n,), dtype=torch.uint8, device=dev)
        s = stream if stream is not None else torch.cuda.current_stream(rays.device).cuda_stream
        check(lib().sortgpu_render_ray_dump(self.handle, ctypes.byref(cam), width, height, spp, depth, seed, pix_begin,
                                            pix_end, rays.data_ptr(), hits.data_ptr(), valid.data_ptr(), s),
              self.accel.ctx.handle)
        return rays, hits, valid

    def li(self, rays, keys, seed=1, max_depth=None, stream=None):
        """Batched Integrator::Li: rays (n,8) and keys (n,2) = (pixel, sample); returns (n,3)."""
        import torch
        dev = f"cuda:{self.accel.ctx.device}"
        r = rays if isinstance(rays, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(rays, np.float32)).to(dev)
        k = keys if isinstance(keys, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(keys, np.uint32).view(np.int32)).to(dev)
        n = r.shape[0]
        out = torch.empty((n, 3), dtype=torch.float32, device=dev)
        s = stream if stream is not None else torch.cuda.current_stream(out.device).cuda_stream
        depth = self.max_recursive_depth if max_depth is None else max_depth
        check(lib().sortgpu_li(self.handle, r.data_ptr(), k.data_ptr(), n, depth, seed, out.data_ptr(), s),
              self.accel.ctx.handle)
        return out

    def stats(self):
        out = np.zeros(4, np.uint64)
        check(lib().sortgpu_render_stats(self.handle, out.ctypes.data), self.accel.ctx.handle)
        return dict(zip(["extension_rays", "shadow_rays", "paths", "gpu_us"], [int(x) for x in out]))

    def set_timing(self, on=True):
        """Stage timing of later renders (events around each bounce's extension traversal,
        shading kernel and shadow rays); read with stage_times()."""
        check(lib().sortgpu_render_set_timing(self.handle, int(bool(on))), self.accel.ctx.handle)

    def stage_times(self):
        """Summed milliseconds per stage of the last timed render and the bounces timed."""
        ms = np.zeros(3, np.float32)
        n = ctypes.c_uint32(0)
        check(lib().sortgpu_render_stage_times(self.handle, ms.ctypes.data, ctypes.byref(n)), self.accel.ctx.handle)
        return {"extend_ms": float(ms[0]), "shade_ms": float(ms[1]), "shadow_ms": float(ms[2]), "bounces": int(n.value)}
