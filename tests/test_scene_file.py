"""The .sort scene loader (sortgpu_scene_file_*) against the oracle's restatement of SORT's reading.

The oracle (oracle/scene_file_oracle.py) is pinned first, independently of its own formulas:
StringIDs against zlib's CRC-32 (SORT's CRC starts from 0 where zlib's starts from ~0, which
zlib.crc32(s, 0xFFFFFFFF) reproduces), the matrix inverse against products with the input and
against exact inverses, transforms, normals, light radiance and the camera's field of view
against closed forms. The loader must then equal the oracle bit for bit on every array of
randomised scenes (all record kinds: cameras, several meshes with arbitrary transforms, every
light type, sky, hair, volumes, shader templates, the material kinds and the integrators'
payloads), and refuse every broken stream with a message naming the byte offset.
"""
import math
import zlib

import numpy as np
import pytest

from oracle import scene_file_oracle as O
from sort_amd import SortGpuError, scene_file, scenes

ARRAYS = ("pos", "idx", "normals", "tri_material", "quad_radiance")


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _within_ulp(a, b, n=1):
    a, b = np.float32(a), np.float32(b)
    return a.shape == b.shape and bool(np.all(np.abs(a - b) <= n * np.spacing(np.maximum(np.abs(a), np.abs(b)))))


def _assert_generated(g, o):
    """Texture coordinates and tangents. Stored coordinates are copied and generated ones come
    from acosf / atan2f of the host's libm in the loader and from double rounded to float in the
    oracle: an ulp of the angle, scaled by 1/pi or 1/2pi, is at most two ulps of the coordinate.
    The tangents are the oracle's smoothing of the loader's own coordinates, bit for bit, and
    equal the oracle's wherever the coordinates are."""
    assert _within_ulp(g["uv"], o["uv"], 2)
    assert _same(g["tangents"], O.smooth_tangents(g["pos"], g["uv"], g["idx"]))
    if _same(g["uv"], o["uv"]):
        assert _same(g["tangents"], o["tangents"])
    assert np.array_equal(np.isnan(g["tangents"]), np.isnan(o["tangents"]))


def _assert_equal_scenes(g, o):
    for k in ARRAYS:
        assert _same(g[k], o[k]), k
    _assert_generated(g, o)
    assert len(g["quads"]) == len(o["quads"])
    for qg, qo in zip(g["quads"], o["quads"]):
        assert _same(np.float32(qg["mat"]), np.float32(qo["mat"])) and _same(np.float32(qg["inv"]), np.float32(qo["inv"]))
        assert np.float32(qg["size_x"]) == np.float32(qo["size_x"]) and np.float32(qg["size_y"]) == np.float32(qo["size_y"])
    assert len(g["materials"]) == len(o["materials"])
    for mg, mo in zip(g["materials"], o["materials"]):
        assert mg["type"] == mo["type"]
        for key in mg:
            assert np.array_equal(np.float32(mg[key]), np.float32(mo[key])), key
    assert g["material_names"] == o["material_names"]
    assert np.array_equal(np.float32(g["env"]), np.float32(o["env"]))
    assert g["skipped"] == o["skipped"]
    for key in o["config"]:
        assert g["config"][key] == pytest.approx(o["config"][key]), key
    if o["camera"] is None:
        assert g["camera"] is None
    else:
        for key in ("eye", "target", "up"):
            assert np.array_equal(np.float32(g["camera"][key]), np.float32(o["camera"][key]))
        # tan / atan: libm float forms in the loader, double rounded to float in the oracle
        assert g["camera"]["fov_y_deg"] == pytest.approx(o["camera"]["fov_y_deg"], rel=1e-6)


# ---- the oracle, pinned ----------------------------------------------------------------------

@pytest.mark.parametrize("name", ["verification bits", "End of Entities", "MeshVisual", "PathTracing", "Obvh",
                                  "SQUARE", "a", ""])
def test_sid_is_crc32_from_zero(name):
    assert O.sid(name) == zlib.crc32(name.encode(), 0xFFFFFFFF)
    assert scene_file.sid(name) == O.sid(name)


def test_inverse_pinned():
    g = np.random.default_rng(3)
    for _ in range(50):
        m = (np.eye(4) + 0.3 * g.normal(size=(4, 4))).astype(np.float32)
        inv = O.inverse(m.reshape(16))
        prod = m.astype(np.float64) @ inv.reshape(4, 4).astype(np.float64)
        assert np.abs(prod - np.eye(4)).max() < 1e-4
    # exact cases: power-of-two scales with small integer translations
    m = np.array([[2, 0, 0, 3], [0, 0.5, 0, -1], [0, 0, 4, 0.5], [0, 0, 0, 1]], np.float32)
    want = np.array([[0.5, 0, 0, -1.5], [0, 2, 0, 2], [0, 0, 0.25, -0.125], [0, 0, 0, 1]], np.float32)
    assert np.array_equal(O.inverse(m.reshape(16)).reshape(4, 4), want)
    assert O.inverse(np.zeros(16, np.float32)) is None
    # the determinant is the product of a triangular matrix's diagonal
    t = np.triu(g.normal(size=(4, 4))).astype(np.float32)
    assert float(O.determinant([np.float32(x) for x in t.reshape(16)])) == pytest.approx(float(np.prod(np.diag(t))), rel=1e-5)


def _quad_mesh(z=0.25, mat=0):
    verts = np.array([[0, 0, z, 0, 0, 2, 0, 0], [1, 0, z, 0, 0, 2, 1, 0], [1, 1, z, 0, 0, 2, 1, 1],
                      [0, 1, z, 0, 0, 2, 0, 1]], np.float32)
    tris = np.array([[0, 1, 2, mat], [0, 2, 3, mat]], np.int32)
    return verts, tris


def test_closed_form_transforms_and_lights():
    verts, tris = _quad_mesh()
    m = np.array([[2, 0, 0, 1], [0, 4, 0, 2], [0, 0, 0.5, 3], [0, 0, 0, 1]], np.float32)
    light_m = np.array([[0, 0, 3, 5], [2, 0, 0, 6], [0, 1, 0, 7], [0, 0, 0, 1]], np.float32)  # columns scaled 2, 1, 3
    data = O.write_scene(materials=[{"name": "white", "kind": O.DIFFUSE, "params": {"Diffuse": (0.5, 0.25, 1.0)}}],
                         meshes=[{"verts": verts, "tris": tris, "matrix": m, "has_uv": True}],
                         lights=[{"type": "AREA", "matrix": light_m, "energy": 10.0, "color": (1.0, 0.5, 0.25),
                                  "shape": "RECTANGLE", "size": 0.5, "size_y": 0.25}])
    for sc in (O.read_scene(data), scene_file.parse(data)):
        want = verts[:, :3] * np.float32([2, 4, 0.5]) + np.float32([1, 2, 3])
        assert np.array_equal(sc["pos"], want)
        # normal (0,0,2) -> normalised (0,0,1) -> transposed inverse: (0, 0, 1/0.5) = (0, 0, 2)
        assert np.array_equal(sc["normals"], np.tile(np.float32([0, 0, 2]), (4, 1)))
        assert np.array_equal(sc["uv"], verts[:, 6:8])
        # uv = (x, y) before the transform: dp/du is the transform's first column, unit length
        assert np.array_equal(sc["tangents"], np.tile(np.float32([1, 0, 0]), (4, 1)))
        assert sc["idx"].tolist() == [[0, 1, 2], [0, 2, 3]] and sc["tri_material"].tolist() == [0, 0]
        q = sc["quads"][0]
        # the light's scale (column lengths 2, 1, 3) moves into the shape: unit columns remain
        assert np.allclose(np.float32(q["mat"]), [[0, 0, 1, 5], [1, 0, 0, 6], [0, 1, 0, 7], [0, 0, 0, 1]], atol=0)
        assert np.allclose(np.float32(q["mat"]) @ np.float32(q["inv"]), np.eye(4), atol=1e-6)
        assert (q["size_x"], q["size_y"]) == (0.5, 0.25)
        scale = 0.8 * 10.0 / (0.5 * 2.0 * 0.25 * 1.0 * 3.1415926)
        assert np.allclose(sc["quad_radiance"][0], np.array([1.0, 0.5, 0.25]) * scale, rtol=1e-6)
        assert sc["materials"][0]["base"] == pytest.approx((0.5, 0.25, 1.0))
        assert sc["env"] == (0.0, 0.0, 0.0)  # an area light: no default sky


def test_camera_field_of_view_closed_forms():
    # vertical field of view for each sensor fit: auto (wide image -> horizontal), horizontal, vertical
    fov = math.radians(60.0)
    for fit, w, h, sensor, want_y in [
        (0, 200, 100, (36.0, 24.0), 2 * math.atan(math.tan(fov / 2) / 2.0)),  # auto, aspect 2 > 1: fov is horizontal
        (1, 100, 200, (36.0, 24.0), 2 * math.atan(math.tan(fov / 2) / 0.5)),  # horizontal fit
        (2, 200, 100, (36.0, 24.0), fov),                                     # vertical fit
        (0, 200, 100, (0.0, 0.0), fov),                                       # no sensor: vertical
    ]:
        data = O.write_scene(camera={"eye": (0, 0, 0), "up": (0, 1, 0), "target": (0, 0, 1), "width": w, "height": h,
                                     "sensor": sensor, "fit": fit, "fov": fov})
        for sc in (O.read_scene(data), scene_file.parse(data)):
            assert math.radians(sc["camera"]["fov_y_deg"]) == pytest.approx(want_y, rel=1e-6), (fit, w, h)
            assert sc["env"] == (0.5, 0.5, 0.5)  # no light at all: SORT's default sky
            assert sc["skipped"] == {}


def test_generated_uv_closed_forms():
    """Mesh::GenUV on an octahedron about the origin: the six axis directions have known
    spherical angles (theta from +y, phi from +x towards +z)."""
    axes = np.float32([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])
    verts = np.zeros((6, 8), np.float32)
    verts[:, :3] = axes
    verts[:, 6:8] = 7.0  # stored coordinates that has_uv = False discards
    tris = np.int32([[0, 2, 4, 0], [4, 2, 1, 0], [1, 2, 5, 0], [5, 2, 0, 0]])
    want = [(0.5, 0.0), (0.5, 0.5), (0.0, 0.0), (1.0, 0.0), (0.5, 0.25), (0.5, 0.75)]
    data = O.write_scene(meshes=[{"verts": verts, "tris": tris, "has_uv": False}])
    for sc in (O.read_scene(data), scene_file.parse(data)):
        assert np.abs(sc["uv"] - np.float32(want)).max() < 1e-6
    stored = scene_file.parse(O.write_scene(meshes=[{"verts": verts, "tris": tris, "has_uv": True}]))
    assert np.all(stored["uv"] == 7.0)


def _tangents(verts, tris, matrix=None):
    mesh = {"verts": np.float32(verts), "tris": np.int32(tris), "has_uv": True}
    if matrix is not None:
        mesh["matrix"] = np.float32(matrix)
    data = O.write_scene(meshes=[mesh])
    o, g = O.read_scene(data), scene_file.parse(data)
    _assert_generated(g, o)
    return g["tangents"]


def _v(p, uv):
    return list(p) + [0.0, 0.0, 1.0] + list(uv)


def test_smooth_tangent_closed_forms():
    """Mesh::genTagentForTri / GenSmoothTagent against closed forms: dp/du of an affine image of
    the uv plane, the sum at a shared vertex, coordinateSystem's axis for degenerate coordinates
    and NaN for a degenerate triangle (a zero normal in coordinateSystem, as in the reference)."""
    g = np.random.default_rng(5)
    # a uv grid placed by an affine transform: every triangle's dp/du is the first column
    xs = np.float32(np.linspace(0.0, 1.0, 4))
    verts = [_v((x, y, 0.0), (x, y)) for y in xs for x in xs]
    tris = [[r * 4 + c, r * 4 + c + 1, r * 4 + c + 5, 0] for r in range(3) for c in range(3)] + \
           [[r * 4 + c, r * 4 + c + 5, r * 4 + c + 4, 0] for r in range(3) for c in range(3)]
    for _ in range(5):
        m = np.eye(4, dtype=np.float32)
        m[:3, :3] = g.normal(size=(3, 3))
        m[:3, 3] = g.normal(size=3)
        col = m[:3, 0] / np.linalg.norm(m[:3, 0])
        t = _tangents(verts, tris, m)
        assert np.abs(t - col).max() < 1e-5
    # a vertex shared by two triangles takes the normalised sum of their tangents
    verts = [_v((0, 0, 0), (0, 0)), _v((1, 0, 0), (1, 0)), _v((0, 1, 0), (0, 1)), _v((0, 0, 2), (1, 0))]
    t = _tangents(verts, [[0, 1, 2, 0], [0, 3, 2, 0]])  # dp/du = (1, 0, 0) and (0, 0, 2)
    s5 = 1 / math.sqrt(5.0)
    assert np.abs(t - np.float32([[s5, 0, 2 * s5], [1, 0, 0], [s5, 0, 2 * s5], [0, 0, 1]])).max() < 1e-6
    # equal texture coordinates: coordinateSystem's first axis about the face normal
    p = np.float32([[0, 0, 0], [1, 0, 0.5], [0, 1, 2]])
    n = np.cross(p[0] - p[1], p[0] - p[2]).astype(np.float64)
    n /= np.linalg.norm(n)
    want = (np.array([-n[2], 0, n[0]]) / math.hypot(n[0], n[2]) if abs(n[0]) > abs(n[1])
            else np.array([0, n[2], -n[1]]) / math.hypot(n[1], n[2]))
    t = _tangents([_v(q, (0.3, 0.3)) for q in p], [[0, 1, 2, 0]])
    assert np.abs(t - want).max() < 1e-6
    # a triangle with two equal corners has no normal: its vertices' tangents are NaN
    t = _tangents([_v((0, 0, 0), (0, 0)), _v((1, 0, 0), (1, 0)), _v((1, 0, 0), (1, 0)), _v((0, 1, 0), (0, 1))],
                  [[0, 1, 2, 0], [0, 1, 3, 0]])
    assert np.isnan(t[:3]).all() and np.abs(t[3] - np.float32([1, 0, 0])).max() < 1e-6


def test_materials_map_to_gpu_records():
    mats = [
        {"name": "lam", "kind": O.DIFFUSE, "params": {"Diffuse": (0.1, 0.2, 0.3)}},
        {"name": "metal", "kind": O.GGX_NODE, "params": {"BaseColor": (0.9, 0.8, 0.7), "InteriorIOR": (0.2, 0.9, 1.1),
                                                         "AbsorptionCoefficient": (3.9, 2.4, 2.1), "RoughnessU": 0.3}},
        {"name": "glass", "kind": O.GLASS, "params": {"Reflectance": (1, 1, 1), "Transmittance": (0.9, 0.95, 1.0),
                                                      "RoughnessU": 0.05}},
        {"name": "disney", "kind": O.DISNEY_NODE, "params": {"BaseColor": (0.8, 0.4, 0.2), "Metallic": 0.25,
                                                             "Roughness": 0.5, "Clearcoat": 0.5}},
        {"name": "cloth", "kind": O.OTHER},                                       # no GPU BXDF
        {"name": "textured", "kind": O.DIFFUSE, "texture_input": True},          # a node drives Diffuse
        {"name": "aniso", "kind": O.GGX_NODE, "params": {"RoughnessU": 0.1, "RoughnessV": 0.4}},
        {"name": "none", "kind": None},                                           # invalid surface shader
        {"name": "tint", "kind": O.TRANSPARENT_NODE, "params": {"Attenuation": (0.5, 0.25, 0.8)}},
        {"name": "lam_tr", "kind": O.DIFFUSE, "transparent": True},              # transparency elsewhere in a graph
    ]
    verts, tris = _quad_mesh()
    tris[0, 3], tris[1, 3] = -1, 42  # -1 and an unknown id: the default material
    data = O.write_scene(materials=mats, meshes=[{"verts": verts, "tris": tris}])
    o, g = O.read_scene(data), scene_file.parse(data)
    _assert_equal_scenes(g, o)
    m = g["materials"]
    assert [x["type"] for x in m] == [0, 1, 2, 3, 0, 0, 1, 0, 4, 0, 0]
    assert m[8]["base"] == pytest.approx((0.5, 0.25, 0.8))
    assert m[0]["base"] == pytest.approx((0.1, 0.2, 0.3))
    assert m[1]["eta"] == pytest.approx((0.2, 0.9, 1.1)) and m[1]["k"] == pytest.approx((3.9, 2.4, 2.1))
    assert m[1]["roughness"] == pytest.approx(0.3)
    assert m[2]["ior"] == 1.5 and m[2]["trans"] == pytest.approx((0.9, 0.95, 1.0))
    assert m[3]["metallic"] == pytest.approx(0.25) and m[3]["clearcoat"] == pytest.approx(0.5)
    assert m[4]["base"] == (1.0, 1.0, 1.0) and m[7]["base"] == (1.0, 1.0, 1.0) and m[10]["base"] == (1.0, 1.0, 1.0)
    assert g["material_names"] == [x["name"] for x in mats] + ["default"]
    assert g["tri_material"].tolist() == [10, 10]
    assert g["skipped"] == {"unmapped_materials": 1, "approximated_materials": 3}


def _random_scene(seed):
    g = np.random.default_rng(seed)
    meshes = []
    for k in range(3):
        nv = int(g.integers(3, 60))
        verts = g.normal(size=(nv, 8)).astype(np.float32)
        verts[0, 3:6] = 0.0  # a zero normal stays zero
        nt = int(g.integers(1, 80))
        tris = np.concatenate([g.integers(0, nv, (nt, 3)), g.integers(-1, 5, (nt, 1))], 1).astype(np.int32)
        m = (np.eye(4) + 0.5 * g.normal(size=(4, 4))).astype(np.float32)
        if k == 1:
            m[3] = [0, 0, 0, 1]
        meshes.append({"verts": verts, "tris": tris, "matrix": m, "has_uv": bool(k % 2),
                       "volume": (2, 3, 4) if k == 2 else None})
    lights = [
        {"type": "AREA", "matrix": (np.eye(4) + 0.3 * g.normal(size=(4, 4))).astype(np.float32), "energy": 20.0,
         "color": (1.0, 0.9, 0.8), "shape": "SQUARE", "size": 0.3},
        {"type": "AREA", "matrix": np.diag([1.5, 2.0, 0.5, 1.0]).astype(np.float32), "energy": 5.0,
         "color": (0.2, 0.3, 0.4), "shape": "RECTANGLE", "size": 1.0, "size_y": 2.0},
        {"type": "AREA", "matrix": np.eye(4, dtype=np.float32), "energy": 5.0, "color": (1, 1, 1), "shape": "DISK",
         "size": 1.0},
        {"type": "POINT", "matrix": np.eye(4, dtype=np.float32), "energy": 100.0, "color": (1, 1, 1)},
        {"type": "SPOT", "matrix": np.eye(4, dtype=np.float32), "energy": 100.0, "color": (1, 1, 1),
         "spot": (20.0, 30.0)},
        {"type": "SUN", "matrix": np.eye(4, dtype=np.float32), "energy": 3.0, "color": (1, 1, 1)},
    ]
    mats = [{"name": f"m{k}", "kind": kind, "params": {}} for k, kind in
            enumerate([O.DIFFUSE, O.GGX_NODE, O.GLASS, O.DISNEY_NODE, O.DIFFUSE])]
    cam = {"eye": (1.0, 2.0, 3.0), "up": (0.0, 0.0, 1.0), "target": (0.0, 0.5, 0.0), "width": 320, "height": 240,
           "sensor": (36.0, 24.0), "fit": 0, "fov": 0.7, "lens": 0.05}
    return dict(materials=mats, camera=cam, meshes=meshes, lights=lights,
                sky={"intensity": 2.0, "color": (0.1, 0.2, 0.3), "image": "sky.hdr"},
                resources=[("tex.png", "Texture2D")], extra_entities={"hair": [[[(0, 0, 0), (0, 0, 1), (0, 1, 1)], [(1, 1, 1), (1, 1, 2)]]]},
                accel=("Qbvh", 20, 8))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_loader_equals_oracle_on_random_scenes(seed):
    data = O.write_scene(**_random_scene(seed))
    o, g = O.read_scene(data), scene_file.parse(data)
    _assert_equal_scenes(g, o)
    assert g["ntris"] == sum(len(m["tris"]) for m in _random_scene(seed)["meshes"])
    assert len(g["quads"]) == 2
    assert g["skipped"] == {"punctual_lights": 3, "disk_lights": 1, "hair": 2, "volumes": 1, "sky_image": 1,
                            "camera_lens_or_pixel_aspect": 1}
    assert np.float32(g["env"]) == pytest.approx(np.float32([0.2, 0.4, 0.6]))
    assert (g["config"]["accelerator"], g["config"]["accel_max_depth"], g["config"]["accel_max_leaf"]) == \
        (O.sid("Qbvh"), 20, 8)


@pytest.mark.parametrize("integrator", ["AmbientOcclusion", "BidirPathTracing", "InstantRadiosity", "LightTracing",
                                        "WhittedRT", "DirectLight", "PathTracing"])
def test_integrator_payloads(integrator):
    verts, tris = _quad_mesh()
    data = O.write_scene(config={"integrator": integrator, "max_depth": 7, "bssrdf": 3},
                         materials=[{"name": "a", "kind": O.DIFFUSE}], meshes=[{"verts": verts, "tris": tris}])
    g = scene_file.parse(data)
    _assert_equal_scenes(g, O.read_scene(data))
    assert g["config"]["integrator"] == O.sid(integrator) and g["config"]["max_recursive_depth"] == 7
    assert g["config"]["max_bssrdf_bounces"] == (3 if integrator == "PathTracing" else 0)
    assert g["ntris"] == 2


def test_cornell_round_trip(tmp_path):
    sc = scenes.cornell_sphere()
    path = tmp_path / "cornell.sort"
    path.write_bytes(O.export_scene_dict(sc))
    g = scene_file.load(path)
    _assert_equal_scenes(g, O.read_scene(path.read_bytes()))
    assert np.array_equal(g["pos"], sc["pos"]) and np.array_equal(g["idx"], sc["idx"])
    assert np.array_equal(g["tri_material"], sc["tri_material"])
    assert np.abs(g["normals"] - sc["normals"]).max() <= 1.2e-7  # renormalised unit normals
    assert np.array_equal(np.float32(g["quads"][0]["mat"]), np.float32(sc["quads"][0]["mat"]))
    assert np.array_equal(np.float32(g["quads"][0]["inv"]), np.float32(sc["quads"][0]["inv"]))
    assert g["quad_radiance"][0] == pytest.approx(sc["quad_radiance"][0], rel=1e-6)
    assert g["camera"]["fov_y_deg"] == pytest.approx(sc["camera"]["fov_y_deg"], rel=1e-6)


def test_broken_streams_are_refused(tmp_path):
    data = O.write_scene(**_random_scene(4))
    # every truncation point is refused with the offset
    for cut in [0, 3, 4, 40, 200, len(data) // 2, len(data) - 5, len(data) - 1]:
        with pytest.raises(SortGpuError, match="byte") as e:
            scene_file.parse(data[:cut])
        assert e.value.code == -6
    bad = bytearray(O.write_scene(meshes=[{"verts": _quad_mesh()[0], "tris": _quad_mesh()[1]}]))
    vb = O.sid("verification bits").to_bytes(4, "little")
    k = bad.find(vb)
    broken = bad.copy()
    broken[k] ^= 1
    with pytest.raises(SortGpuError, match="not a scene stream"):
        scene_file.parse(bytes(broken))
    # a vertex index out of range
    verts, tris = _quad_mesh()
    tris[1, 2] = 4
    with pytest.raises(SortGpuError, match="vertex index"):
        scene_file.parse(O.write_scene(meshes=[{"verts": verts, "tris": tris}]))
    # an unknown entity class right after the verification bits
    unk = bytearray(bad)
    unk[k + 4:k + 8] = O.sid("TeapotEntity").to_bytes(4, "little")
    with pytest.raises(SortGpuError, match="unknown entity"):
        scene_file.parse(bytes(unk))
    with pytest.raises(SortGpuError, match="version"):
        scene_file.parse(b"\x01\x00\x00\x00" + bytes(data[4:]))
    with pytest.raises(SortGpuError, match="cannot open"):
        scene_file.load(tmp_path / "missing.sort")
    # the oracle refuses the same streams
    with pytest.raises(ValueError):
        O.read_scene(data[: len(data) // 2])


def test_corrupted_streams_parse_or_refuse():
    # seeded corruption (bit flips, saturated bytes, NaN / huge words, truncation): every stream
    # either parses or is refused with the library's error, never anything else
    import random
    rng = random.Random(20240611)
    bases = [O.write_scene(**_random_scene(s)) for s in range(2)]
    outcomes = {"parsed": 0, "refused": 0}
    for _ in range(400):
        b = bytearray(rng.choice(bases))
        if rng.random() < 0.1:
            b = b[:rng.randrange(len(b))]
        for _ in range(rng.choice([1, 1, 2, 4])):
            if not b:
                break
            k = rng.randrange(len(b))
            mode = rng.random()
            if mode < 0.5:
                b[k] ^= 1 << rng.randrange(8)
            elif mode < 0.8:
                b[k] = rng.choice([0, 0xFF, 0x7F, 0x80])
            else:
                k4 = k - k % 4
                b[k4:k4 + 4] = rng.choice([b"\xff\xff\xff\xff", b"\x00\x00\x00\x80", b"\xff\xff\xff\x7f",
                                           b"\x00\x00\xc0\x7f"])
        try:
            scene_file.parse(bytes(b))
            outcomes["parsed"] += 1
        except SortGpuError as e:
            assert e.code == -6, e
            outcomes["refused"] += 1
    assert outcomes["parsed"] > 50 and outcomes["refused"] > 50, outcomes
