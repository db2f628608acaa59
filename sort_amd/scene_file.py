"""SORT scene files (.sort) over the C ABI: ``load(path)`` / ``parse(data)`` read the stream that
SORT's Blender exporter writes (blender-plugin/addons/sortblend/exporter.py:181-440) the way the
renderer reads it (ImageEvaluation::loadConfig, work/image_evaluation/image_evaluation.cpp:354-378;
MatManager::ParseMatFile, material/matmanager.cpp:50-290; Scene::LoadScene, core/scene.cpp:112-180)
and return the scene as the dictionary the accelerator and the path tracer take (the same keys as
``scenes.cornell_sphere()``), plus the file's configuration.

The reading is native (``sortgpu_scene_file_*`` in libsortgpu.so); this module only copies the
arrays out.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

from . import _lib
from ._lib import check, lib

# bits of ``skipped`` (SORTGPU_FILE_* in sortgpu.h)
SKIPPED = ("punctual_lights", "disk_lights", "hair", "volumes", "sky_image", "unmapped_materials",
           "approximated_materials", "camera_lens_or_pixel_aspect")


def sid(name: str) -> int:
    """SORT's StringID of a name (core/strid.h:39-60): bitwise CRC-32, polynomial 0xEDB88320,
    initial value 0, final complement."""
    crc = 0
    for ch in name.encode("ascii"):
        crc ^= ch
        for _ in range(8):
            crc = (crc >> 1) ^ (0xEDB88320 if crc & 1 else 0)
    return (~crc) & 0xFFFFFFFF


class SceneFile:
    """An open scene file; ``scene()`` copies it into numpy arrays."""

    def __init__(self, handle):
        self.handle = handle

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "handle", None):
            lib().sortgpu_scene_file_destroy(self.handle)
            self.handle = None

    def info(self) -> _lib.SceneFileInfo:
        return lib().sortgpu_scene_file_get(self.handle).contents

    def scene(self) -> dict:
        i = self.info()

        def arr(ptr, n, dt, shape):
            if n == 0:
                return np.zeros(shape, dt)
            buf = (ctypes.c_char * (n * np.dtype(dt).itemsize)).from_address(ptr)
            return np.frombuffer(buf, dt).copy().reshape(shape)

        pos = arr(i.pos, i.nverts * 3, np.float32, (i.nverts, 3))
        normals = arr(i.normals, i.nverts * 3, np.float32, (i.nverts, 3))
        uv = arr(i.uv, i.nverts * 2, np.float32, (i.nverts, 2))
        tangents = arr(i.tangents, i.nverts * 3, np.float32, (i.nverts, 3))
        idx = arr(i.idx, i.ntris * 3, np.uint32, (i.ntris, 3))
        tri_material = arr(i.tri_material, i.ntris, np.uint32, (i.ntris,))
        qa = arr(i.quads, i.nquads * 34, np.float32, (i.nquads, 34))
        quads = [{"mat": q[:16].reshape(4, 4).copy(), "inv": q[16:32].reshape(4, 4).copy(),
                  "size_x": float(q[32]), "size_y": float(q[33])} for q in qa]
        quad_radiance = arr(i.quad_radiance, i.nquads * 3, np.float32, (i.nquads, 3))
        mats = (_lib.Material * i.nmaterials).from_address(i.materials) if i.nmaterials else []
        materials = [dict(type=m.type, base=tuple(m.base), trans=tuple(m.trans), roughness=m.roughness, ior=m.ior,
                          eta=tuple(m.eta), k=tuple(m.k), metallic=m.metallic, specular=m.specular,
                          specular_tint=m.specular_tint, sheen=m.sheen, sheen_tint=m.sheen_tint,
                          clearcoat=m.clearcoat, clearcoat_gloss=m.clearcoat_gloss) for m in mats]
        names = [lib().sortgpu_scene_file_material_name(self.handle, k).decode() for k in range(i.nmaterials)]
        skipped = {name: int(i.counts[b]) for b, name in enumerate(SKIPPED) if i.skipped >> b & 1}
        camera = None
        if i.has_camera:
            camera = dict(eye=tuple(i.eye), target=tuple(i.target), up=tuple(i.up),
                          fov_y_deg=math.degrees(i.camera_fov_y))
        return {
            "pos": pos, "idx": idx, "quads": quads, "normals": normals, "uv": uv, "tangents": tangents,
            "tri_material": tri_material,
            "materials": materials, "material_names": names, "quad_radiance": quad_radiance,
            "env": tuple(i.env_radiance), "ntris": int(i.ntris), "camera": camera, "clamp": float(i.clamping),
            "config": {"threads": i.threads, "spp": i.spp, "width": i.image_width, "height": i.image_height,
                       "clamping": i.clamping, "integrator": i.integrator,
                       "max_recursive_depth": i.max_recursive_depth, "max_bssrdf_bounces": i.max_bssrdf_bounces,
                       "accelerator": i.accelerator, "accel_max_depth": i.accel_max_depth,
                       "accel_max_leaf": i.accel_max_leaf},
            "camera_record": {"eye": tuple(i.eye), "up": tuple(i.up), "target": tuple(i.target),
                              "width": i.camera_width, "height": i.camera_height, "lens_radius": i.lens_radius,
                              "sensor": (i.sensor_w, i.sensor_h), "aspect_fit": i.aspect_fit,
                              "pixel_aspect": (i.pixel_aspect_w, i.pixel_aspect_h), "fov": i.fov,
                              "fov_y": i.camera_fov_y},
            "skipped": skipped,
        }


def open_file(path: str | os.PathLike) -> SceneFile:
    h = _lib._P()
    check(lib().sortgpu_scene_file_load(os.fsencode(path), ctypes.byref(h)))
    return SceneFile(h)


def open_bytes(data: bytes) -> SceneFile:
    buf = ctypes.create_string_buffer(bytes(data), len(data))
    h = _lib._P()
    check(lib().sortgpu_scene_file_parse(buf, len(data), ctypes.byref(h)))
    return SceneFile(h)


def load(path: str | os.PathLike) -> dict:
    """The scene dictionary of a .sort file (raises SortGpuError for an unreadable file)."""
    f = open_file(path)
    try:
        return f.scene()
    finally:
        f.close()


def parse(data: bytes) -> dict:
    f = open_bytes(data)
    try:
        return f.scene()
    finally:
        f.close()
