"""The library against the committed fixtures of tests/golden (made by tests/golden/make_golden.py):
the scene loader must reproduce the recorded digests of every array of scene_small.sort, and the
EXR encoder the recorded bytes of film_37x21.exr. The oracle must reproduce them too, so that a
change to either side shows up as a mismatch with fixed bytes."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import film_oracle as FO
from oracle import scene_file_oracle as SF
from sort_amd import film, scene_file

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("reader", ["library", "oracle"])
def test_scene_small_fixture(reader):
    data = open(os.path.join(HERE, "scene_small.sort"), "rb").read()
    want = json.load(open(os.path.join(HERE, "scene_small.json")))
    sc = scene_file.parse(data) if reader == "library" else SF.read_scene(data)
    for k in ("pos", "idx", "normals", "tri_material", "quad_radiance"):
        assert list(np.asarray(sc[k]).shape) == want[k]["shape"], k
        assert _digest(sc[k]) == want[k]["sha256"], k
    # one mesh stores texture coordinates, the other has them generated (acos / atan2 of the
    # platform's libm): within two ulps; the tangents smoothed from them within 1e-5
    uv = np.float32(want["uv_values"])
    assert sc["uv"].shape == uv.shape
    assert np.all(np.abs(sc["uv"] - uv) <= 2 * np.spacing(np.maximum(np.abs(sc["uv"]), np.abs(uv))))
    tan = np.array([[np.nan if x is None else x for x in row] for row in want["tangent_values"]], np.float32)
    assert sc["tangents"].shape == tan.shape
    assert np.array_equal(np.isnan(sc["tangents"]), np.isnan(tan))
    ok = ~np.isnan(tan)
    assert np.abs(sc["tangents"][ok] - tan[ok]).max() < 1e-5
    assert len(sc["quads"]) == len(want["quads"])
    for q, w in zip(sc["quads"], want["quads"]):
        assert _digest(np.float32(q["mat"])) == w["mat"] and _digest(np.float32(q["inv"])) == w["inv"]
        assert [q["size_x"], q["size_y"]] == pytest.approx(w["size"])
    for m, w in zip(sc["materials"], want["materials"]):
        for key, v in w.items():
            assert np.allclose(np.float32(m[key]), np.float32(v), rtol=0, atol=0), key
    assert sc["material_names"] == want["material_names"]
    assert np.array_equal(np.float32(sc["env"]), np.float32(want["env"]))
    assert sc["camera"]["fov_y_deg"] == pytest.approx(want["camera_fov_y_deg"], rel=1e-6)
    assert sc["skipped"] == want["skipped"]
    for key, v in want["config"].items():
        assert sc["config"][key] == pytest.approx(v), key
    assert sc["config"]["clamping"] == pytest.approx(10.0)


def test_film_fixture():
    f = np.load(os.path.join(HERE, "film_37x21.npy"))
    want = open(os.path.join(HERE, "film_37x21.exr"), "rb").read()
    assert film.encode_exr(f) == want
    assert FO.exr_file(f, 37, 21) == want
    r = FO.read_exr(want)
    assert (r["width"], r["height"]) == (37, 21)
