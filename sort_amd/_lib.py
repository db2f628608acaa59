"""ctypes binding of libsortgpu.so (the C ABI declared in include/sortgpu.h).

The library is loaded from this package directory. A missing library is an error: there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsortgpu.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "sortgpu.h")

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_OOM = -3
ERR_BUDGET = -4
ERR_STATE = -5

# error flags the kernels raise (taken by Obvh.check_errors / Pool.check_errors)
FLAG_STACK_OVERFLOW = 1
FLAG_BAD_SCENE = 2
FLAG_ATTEN_CAP = 4


class SortGpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sortgpu error {code}: {msg}")
        self.code = code


class Quad(ctypes.Structure):
    _fields_ = [("mat", ctypes.c_float * 16), ("inv", ctypes.c_float * 16),
                ("size_x", ctypes.c_float), ("size_y", ctypes.c_float)]


class Surface(ctypes.Structure):
    _fields_ = [("intersect", ctypes.c_float * 3), ("t", ctypes.c_float), ("gnormal", ctypes.c_float * 3),
                ("normal", ctypes.c_float * 3), ("tangent", ctypes.c_float * 3), ("view", ctypes.c_float * 3),
                ("u", ctypes.c_float), ("v", ctypes.c_float), ("prim", ctypes.c_uint32), ("is_quad", ctypes.c_uint32)]


class BuildOpts(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("max_depth", ctypes.c_uint32),
                ("max_leaf", ctypes.c_uint32), ("build_threads", ctypes.c_uint32)]


class Tuning(ctypes.Structure):
    _fields_ = [("refill", ctypes.c_uint32), ("lds_entries", ctypes.c_uint32),
                ("blocks_per_cu", ctypes.c_uint32), ("leaf_batch", ctypes.c_uint32), ("leaf_min", ctypes.c_uint32),
                ("reorder", ctypes.c_uint32), ("shade_order", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class Material(ctypes.Structure):
    _fields_ = [("type", ctypes.c_uint32), ("base", ctypes.c_float * 3), ("trans", ctypes.c_float * 3),
                ("roughness", ctypes.c_float), ("ior", ctypes.c_float), ("eta", ctypes.c_float * 3),
                ("k", ctypes.c_float * 3), ("metallic", ctypes.c_float), ("specular", ctypes.c_float),
                ("specular_tint", ctypes.c_float), ("sheen", ctypes.c_float), ("sheen_tint", ctypes.c_float),
                ("clearcoat", ctypes.c_float), ("clearcoat_gloss", ctypes.c_float)]


class SceneDesc(ctypes.Structure):
    _fields_ = [("normals", ctypes.c_void_p), ("tangents", ctypes.c_void_p), ("tri_material", ctypes.c_void_p),
                ("materials", ctypes.c_void_p), ("nmaterials", ctypes.c_uint32), ("quad_radiance", ctypes.c_void_p),
                ("env_radiance", ctypes.c_float * 3), ("clamp", ctypes.c_float)]


class Camera(ctypes.Structure):
    _fields_ = [("eye", ctypes.c_float * 3), ("target", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
                ("fov_y", ctypes.c_float)]


class SceneFileInfo(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_uint32), ("spp", ctypes.c_uint32), ("image_width", ctypes.c_uint32),
                ("image_height", ctypes.c_uint32), ("clamping", ctypes.c_float), ("integrator", ctypes.c_uint32),
                ("max_recursive_depth", ctypes.c_int32), ("max_bssrdf_bounces", ctypes.c_int32),
                ("eye", ctypes.c_float * 3), ("up", ctypes.c_float * 3), ("target", ctypes.c_float * 3),
                ("camera_width", ctypes.c_uint32), ("camera_height", ctypes.c_uint32),
                ("lens_radius", ctypes.c_float), ("sensor_w", ctypes.c_float), ("sensor_h", ctypes.c_float),
                ("aspect_fit", ctypes.c_int32), ("pixel_aspect_w", ctypes.c_float), ("pixel_aspect_h", ctypes.c_float),
                ("fov", ctypes.c_float), ("camera_fov_y", ctypes.c_float), ("has_camera", ctypes.c_uint32),
                ("nverts", ctypes.c_uint32), ("ntris", ctypes.c_uint32), ("nquads", ctypes.c_uint32),
                ("nmaterials", ctypes.c_uint32), ("pos", ctypes.c_void_p), ("normals", ctypes.c_void_p),
                ("uv", ctypes.c_void_p), ("tangents", ctypes.c_void_p), ("idx", ctypes.c_void_p),
                ("tri_material", ctypes.c_void_p),
                ("quads", ctypes.c_void_p), ("quad_radiance", ctypes.c_void_p), ("materials", ctypes.c_void_p),
                ("env_radiance", ctypes.c_float * 3), ("accelerator", ctypes.c_uint32),
                ("accel_max_depth", ctypes.c_int32), ("accel_max_leaf", ctypes.c_int32),
                ("skipped", ctypes.c_uint32), ("counts", ctypes.c_uint32 * 8)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("wave_paths", ctypes.c_uint32), ("max_active", ctypes.c_uint32),
                ("inflight_waves", ctypes.c_uint32), ("manual", ctypes.c_uint32), ("host_device", ctypes.c_uint32),
                ("fail_alloc_at", ctypes.c_uint32), ("graphs", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class Request(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_uint32), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("spp", ctypes.c_uint32), ("max_depth", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("seed", ctypes.c_uint64), ("camera", Camera), ("timeout_s", ctypes.c_double)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("status", ctypes.c_int32), ("samples_done", ctypes.c_uint32),
                ("spp", ctypes.c_uint32), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double)]


_lib = None

_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_I = ctypes.c_int

_SIGS = {
    "sortgpu_ctx_create": (_I, [_I, ctypes.POINTER(_P)]),
    "sortgpu_ctx_destroy": (None, [_P]),
    "sortgpu_last_error": (ctypes.c_char_p, [_P]),
    "sortgpu_accel_build": (_I, [_P, _P, _U32, _P, _U32, _P, _U32, ctypes.POINTER(BuildOpts), ctypes.POINTER(_P)]),
    "sortgpu_accel_destroy": (None, [_P]),
    "sortgpu_accel_valid": (_I, [_P]),
    "sortgpu_accel_bbox": (_I, [_P, _P]),
    "sortgpu_accel_info": (_I, [_P, _P]),
    "sortgpu_accel_prim_order": (_I, [_P, _P, ctypes.c_uint64]),
    "sortgpu_accel_tree_dump": (ctypes.c_int64, [_P, _P, ctypes.c_uint64]),
    "sortgpu_trace_closest": (_I, [_P, _P, _U32, _P, _P]),
    "sortgpu_trace_closest_noalloc": (_I, [_P, _P, _U32, _P, _P]),
    "sortgpu_trace_closest1": (_I, [_P, _P, _P]),
    "sortgpu_trace_any1": (_I, [_P, _P, ctypes.POINTER(_I)]),
    "sortgpu_hit_surface": (_I, [_P, _P, _P, _P, _P, _P, ctypes.POINTER(Surface)]),
    "sortgpu_trace_any": (_I, [_P, _P, _U32, _P, _P]),
    "sortgpu_trace_attenuation": (_I, [_P, _P, _U32, _P, _U32, _P, _P]),
    "sortgpu_trace_multi": (_I, [_P, _P, _U32, _P, _P, _U32, _P, _P, _P]),
    "sortgpu_trace_counters": (_I, [_P, _P, _U32, _I, _P, _P, _P, _P]),
    "sortgpu_accel_errors": (_I, [_P, _P, ctypes.POINTER(_U32)]),
    "sortgpu_pool_errors": (_I, [_P, _P, ctypes.POINTER(_U32)]),
    "sortgpu_ctx_set_tuning": (_I, [_P, ctypes.POINTER(Tuning)]),
    "sortgpu_pool_create": (_I, [_P, _P, _U32, ctypes.POINTER(_P)]),
    "sortgpu_pool_destroy": (None, [_P]),
    "sortgpu_pool_info": (_I, [_P, _P]),
    "sortgpu_pool_trace": (_I, [_P, _P, _P, _U32, _I, _P, _P]),
    "sortgpu_scene_create": (_I, [_P, ctypes.POINTER(SceneDesc), ctypes.POINTER(_P)]),
    "sortgpu_scene_destroy": (None, [_P]),
    "sortgpu_render": (_I, [_P, ctypes.POINTER(Camera), _U32, _U32, _U32, _U32, ctypes.c_uint64, _P, _P]),
    "sortgpu_li": (_I, [_P, _P, _P, _U32, _U32, ctypes.c_uint64, _P, _P]),
    "sortgpu_render_ray_dump": (_I, [_P, ctypes.POINTER(Camera), _U32, _U32, _U32, _U32, ctypes.c_uint64, _U32, _U32,
                                     _P, _P, _P, _P]),
    "sortgpu_render_stats": (_I, [_P, _P]),
    "sortgpu_render_set_timing": (_I, [_P, ctypes.c_int]),
    "sortgpu_render_stage_times": (_I, [_P, _P, _P]),
    "sortgpu_server_create": (_I, [_P, ctypes.POINTER(ServerOpts), ctypes.POINTER(_P)]),
    "sortgpu_server_destroy": (None, [_P]),
    "sortgpu_server_add_scene": (_I, [_P, _P, ctypes.POINTER(SceneDesc), ctypes.POINTER(_U32)]),
    "sortgpu_server_submit": (_I, [_P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
    "sortgpu_server_cancel": (_I, [_P, ctypes.c_uint64]),
    "sortgpu_server_info": (_I, [_P, ctypes.c_uint64, ctypes.POINTER(RequestInfo)]),
    "sortgpu_server_wait": (_I, [_P, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(RequestInfo)]),
    "sortgpu_server_result": (_I, [_P, ctypes.c_uint64, _P, ctypes.c_uint64]),
    "sortgpu_server_message": (_I, [_P, ctypes.c_uint64, ctypes.c_char_p, _U32]),
    "sortgpu_server_release": (_I, [_P, ctypes.c_uint64]),
    "sortgpu_server_step": (_I, [_P, ctypes.POINTER(_I)]),
    "sortgpu_server_drain": (_I, [_P, ctypes.c_double]),
    "sortgpu_server_stats": (_I, [_P, _P]),
    "sortgpu_scene_file_load": (_I, [ctypes.c_char_p, ctypes.POINTER(_P)]),
    "sortgpu_scene_file_parse": (_I, [_P, ctypes.c_uint64, ctypes.POINTER(_P)]),
    "sortgpu_scene_file_get": (ctypes.POINTER(SceneFileInfo), [_P]),
    "sortgpu_scene_file_material_name": (ctypes.c_char_p, [_P, _U32]),
    "sortgpu_scene_file_destroy": (None, [_P]),
    "sortgpu_exr_size": (ctypes.c_uint64, [_U32, _U32]),
    "sortgpu_film_encode_exr": (_I, [_P, _I, _U32, _U32, _P, ctypes.c_uint64, _P]),
    "sortgpu_film_write_exr": (_I, [_P, _I, _U32, _U32, ctypes.c_char_p, _P]),
    "sortgpu_display_packet": (ctypes.c_int64, [_I, ctypes.c_char_p, _P, _U32, _U32, _U32, _U32, _U32, _U32, _P,
                                                ctypes.c_uint64]),
    "sortgpu_display_connect": (_I, [ctypes.c_char_p, ctypes.c_uint16, ctypes.POINTER(_P)]),
    "sortgpu_display_send_film": (_I, [_P, _I, ctypes.c_char_p, _P, _I, _U32, _U32, _U32, _P]),
    "sortgpu_display_close": (None, [_P]),
}


def header_symbols(header: str = HEADER) -> list[str]:
    """Every function a header declares (default: the public header)."""
    text = open(header).read()
    return sorted(set(re.findall(r"\b(sortgpu_[a-z0-9_]+)\s*\(", text)))


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        # torch ships its own HIP runtime under the same soname as the system one; whichever loads
        # first serves both. Loading torch first keeps the process on the runtime torch was built
        # for (torch fails to see the GPU when it finds the other one already loaded).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def check(rc: int, ctx=None) -> int:
    if rc < 0:
        msg = lib().sortgpu_last_error(ctx)
        raise SortGpuError(rc, msg.decode() if msg else "")
    return rc
