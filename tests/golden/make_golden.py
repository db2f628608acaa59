"""Writes the committed fixtures of tests/test_golden.py (run from the repository root):

* scene_small.sort — a .sort stream in the exporter's layout (two meshes with transforms, square,
  rectangle, disk and point lights, a sky, four material kinds, a camera with a sensor fit),
  written by the oracle's restatement of the exporter;
* scene_small.json — what the renderer's reading of it gives (the oracle's restatement): array
  shapes and SHA-256 digests of their bytes (the values of the texture coordinates and tangents),
  materials, environment, camera, skipped content;
* film_37x21.npy / film_37x21.exr — a seeded film holding special values and the EXR file SORT's
  output path (tinyexr's SaveEXR) writes for it.

The reference ships no scene files or images; these fixtures freeze the restatements' outputs so
that the library is checked against fixed bytes rather than against code that could drift with it.
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import film_oracle as FO  # noqa: E402
from oracle import scene_file_oracle as SF  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def scene_small():
    g = np.random.default_rng(20261202)
    meshes = []
    for k in range(2):
        verts = g.normal(size=(12, 8)).astype(np.float32)
        tris = np.concatenate([g.integers(0, 12, (10, 3)), g.integers(-1, 4, (10, 1))], 1).astype(np.int32)
        m = (np.eye(4) + 0.25 * g.normal(size=(4, 4))).astype(np.float32)
        m[3] = [0, 0, 0, 1]
        meshes.append({"verts": verts, "tris": tris, "matrix": m, "has_uv": k == 0})
    mats = [{"name": "white", "kind": SF.DIFFUSE, "params": {"Diffuse": (0.7, 0.7, 0.7)}},
            {"name": "copper", "kind": SF.GGX_NODE, "params": {"BaseColor": (0.95, 0.64, 0.54), "RoughnessU": 0.25}},
            {"name": "glass", "kind": SF.GLASS, "params": {"RoughnessU": 0.05}},
            {"name": "paint", "kind": SF.DISNEY_NODE, "params": {"BaseColor": (0.2, 0.3, 0.8), "Metallic": 0.0,
                                                                  "Clearcoat": 0.5}}]
    lights = [{"type": "AREA", "matrix": np.array([[1, 0, 0, 0.5], [0, -1, 0, 2.0], [0, 0, -1, 0.5], [0, 0, 0, 1]],
                                                   np.float32), "energy": 30.0, "color": (1.0, 0.9, 0.8),
               "shape": "SQUARE", "size": 0.5},
              {"type": "AREA", "matrix": np.diag([2.0, 1.0, 0.5, 1.0]).astype(np.float32), "energy": 8.0,
               "color": (0.3, 0.4, 1.0), "shape": "RECTANGLE", "size": 0.4, "size_y": 0.2},
              {"type": "AREA", "matrix": np.eye(4, dtype=np.float32), "energy": 1.0, "color": (1, 1, 1),
               "shape": "DISK", "size": 0.3},
              {"type": "POINT", "matrix": np.eye(4, dtype=np.float32), "energy": 50.0, "color": (1, 1, 1)}]
    cam = {"eye": (0.5, 0.5, -2.0), "up": (0.0, 1.0, 0.0), "target": (0.5, 0.5, 0.5), "width": 320, "height": 180,
           "sensor": (36.0, 24.0), "fit": 0, "fov": 0.6}
    return SF.write_scene(config={"spp": 32, "width": 320, "height": 180, "max_depth": 6, "clamping": 10.0},
                          materials=mats, camera=cam, meshes=meshes, lights=lights,
                          sky={"intensity": 0.5, "color": (0.2, 0.3, 0.4)}, accel=("Obvh", 24, 8))


def main():
    data = scene_small()
    with open(os.path.join(HERE, "scene_small.sort"), "wb") as f:
        f.write(data)
    sc = SF.read_scene(data)
    expect = {k: {"shape": list(np.asarray(sc[k]).shape), "sha256": digest(sc[k])}
              for k in ("pos", "idx", "normals", "tri_material", "quad_radiance")}
    # generated texture coordinates go through acos / atan2, whose last bit depends on the libm:
    # their values (and the tangents smoothed from them) are recorded instead of digests
    expect["uv_values"] = sc["uv"].tolist()
    expect["tangent_values"] = [[None if np.isnan(x) else float(x) for x in row] for row in sc["tangents"]]
    expect["quads"] = [{"mat": digest(np.float32(q["mat"])), "inv": digest(np.float32(q["inv"])),
                        "size": [q["size_x"], q["size_y"]]} for q in sc["quads"]]
    expect["materials"] = sc["materials"]
    expect["material_names"] = sc["material_names"]
    expect["env"] = list(sc["env"])
    expect["camera_fov_y_deg"] = sc["camera"]["fov_y_deg"]
    expect["skipped"] = sc["skipped"]
    expect["config"] = {k: (float(v) if isinstance(v, (float, np.floating)) else int(v)) for k, v in sc["config"].items()}
    with open(os.path.join(HERE, "scene_small.json"), "w") as f:
        json.dump(expect, f, indent=1, sort_keys=True, default=float)
    g = np.random.default_rng(7)
    film = (g.random((21, 37, 3)) * 3.0).astype(np.float32)
    film[0, 0] = [np.inf, 65520.0, 1e-40]
    film[0, 1] = [1.0 + 2.0 ** -11, 2.0 ** -15 + 2.0 ** -25, -(1.0 + 2.0 ** -11)]
    np.save(os.path.join(HERE, "film_37x21.npy"), film)
    with open(os.path.join(HERE, "film_37x21.exr"), "wb") as f:
        f.write(FO.exr_file(film, 37, 21))


if __name__ == "__main__":
    main()
