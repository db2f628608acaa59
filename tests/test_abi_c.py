"""The drop-in boundary from a plain C caller (no GPU needed).

tests/native/abi_layout.c includes include/sortgpu.h as C, is built with gcc against
libsortgpu.so, prints the size and field offsets of every public struct and drives the host-side
entry points on a host-only context. The Python binding's ctypes layouts (sort_amd/_lib.py) must
equal the C compiler's, field by field, and the C caller must see the same trees as the binding.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from sort_amd import Obvh, Qbvh
from sort_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# C struct -> ctypes class of the binding (rays and hits travel as (n, 8) / (n, 4) float32 arrays)
CTYPES = {
    "sortgpu_quad": _lib.Quad, "sortgpu_build_opts": _lib.BuildOpts, "sortgpu_tuning": _lib.Tuning,
    "sortgpu_material": _lib.Material, "sortgpu_scene_desc": _lib.SceneDesc, "sortgpu_camera": _lib.Camera,
    "sortgpu_server_opts": _lib.ServerOpts, "sortgpu_request": _lib.Request,
    "sortgpu_request_info": _lib.RequestInfo, "sortgpu_scene_file_info": _lib.SceneFileInfo,
    "sortgpu_surface": _lib.Surface,
}


@pytest.fixture(scope="module")
def c_output(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    exe = str(tmp_path_factory.mktemp("abi") / "abi_layout")
    libdir = os.path.dirname(_lib.LIB_PATH)
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "native", "abi_layout.c"), "-o", exe, f"-L{libdir}", "-lsortgpu",
           f"-Wl,-rpath,{libdir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    sizes, fields, res = {}, {}, {}
    for ln in r.stdout.splitlines():
        k, *rest = ln.split()
        if k == "S":
            sizes[rest[0]] = int(rest[1])
        elif k == "F":
            s, f = rest[0].split(".")
            fields.setdefault(s, {})[f] = (int(rest[1]), int(rest[2]))
        elif k == "R":
            res[rest[0]] = rest[1:]
    return sizes, fields, res


def test_ctypes_layouts_equal_the_c_compilers(c_output):
    sizes, fields, _ = c_output
    assert sizes["sortgpu_ray"] == 32 and sizes["sortgpu_hit"] == 16
    assert [fields["sortgpu_ray"][f][0] for f in ["ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax"]] == \
        list(range(0, 32, 4))
    assert [fields["sortgpu_hit"][f][0] for f in ["t", "prim", "u", "v"]] == [0, 4, 8, 12]
    for cname, cls in CTYPES.items():
        assert ctypes.sizeof(cls) == sizes[cname], cname
        names = [f[0] for f in cls._fields_]
        assert sorted(names) == sorted(fields[cname]), cname
        for name in names:
            off, size = fields[cname][name]
            d = getattr(cls, name)
            assert (d.offset, d.size) == (off, size), f"{cname}.{name}: ctypes {(d.offset, d.size)} C {(off, size)}"


def _c_scene():
    """The scene abi_layout.c builds: a 2x2x2 grid of small triangles and one quad light."""
    pos, idx, t = [], [], 0
    for z in range(2):
        for y in range(2):
            for x in range(2):
                for v in range(3):
                    pos.append([x + (0.25 if v == 1 else 0.0), y + (0.25 if v == 2 else 0.0), z])
                    idx.append(3 * t + v)
                t += 1
    mat = np.eye(4, dtype=np.float32)
    inv = np.eye(4, dtype=np.float32)
    mat[1, 3], inv[1, 3] = 3.0, -3.0
    quad = {"mat": mat, "inv": inv, "size_x": 0.5, "size_y": 0.5}
    return np.array(pos, np.float32), np.array(idx, np.uint32).reshape(-1, 3), [quad]


@pytest.mark.parametrize("width", [8, 4])
def test_c_caller_sees_the_bindings_trees(c_output, host_ctx, width):
    _, _, res = c_output
    assert res["null_ctx"] == ["-1"] and res["ctx"] == ["0"]
    assert res["null_accel_build"] == ["-1"] and res["bad_width"] == ["-1"]
    assert res[f"build{width}"] == ["0"]
    pos, idx, quads = _c_scene()
    acc = (Obvh if width == 8 else Qbvh)(host_ctx, max_pri_in_leaf=2)
    acc.build(pos, idx, quads)
    info = acc.info()
    got = [int(x) for x in res[f"info{width}"]]
    assert got[0] == 0 and got[1] == width and got[5] == info["prims"] == 9
    assert got[2:7] == [info[k] for k in ["depth", "inner_nodes", "leaves", "prims", "max_stack"]]
    assert [int(x) for x in res[f"order{width}"][1:]] == acc.prim_order().tolist()
    assert [int(x) for x in res[f"dump{width}"][1:]] == acc.tree_dump().tolist()
    assert np.array_equal(np.array(res[f"bbox{width}"][1:], np.float32), acc.bbox())
    assert res[f"query{width}"] == ["-5"]          # host-only context: SORTGPU_ERR_STATE
    assert res[f"flags{width}"] == ["0", "0"]      # nothing raised, flags cleared to 0
    assert res[f"valid{width}"] == ["1"]
