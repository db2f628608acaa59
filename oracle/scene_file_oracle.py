"""Test infrastructure: CPU restatement of SORT's .sort scene stream, both directions.

* ``SortFileWriter`` / ``write_scene``: the byte layout that SORT's Blender exporter writes
  (blender-plugin/addons/sortblend/exporter.py: export_global_config :424-447, materials :726-935,
  export_scene :227-408, export_mesh :488-585; stream/stream.py for the value encodings), driven
  by plain data instead of Blender objects, to make scene files for the tests (the reference ships
  none).
* ``read_scene``: the renderer's reading of such a stream (ImageEvaluation::loadConfig,
  work/image_evaluation/image_evaluation.cpp:354-378; MatManager::ParseMatFile,
  material/matmanager.cpp:50-290; Material::Serialize, material/material.cpp:153-228;
  Scene::LoadScene, core/scene.cpp:112-180, with the entity readers of entity/*.cpp and
  core/mesh.cpp:26-40, 111-180) flattened the way the C ABI loader does: world-space float32
  positions and normals in the reference's operation order, quads for square and rectangular
  area lights, the GPU material each shader graph maps to.

Only tests/ may use this module. Pinned by tests/test_scene_file.py: StringIDs against zlib's
CRC-32, the inverse against exact and near-identity products, closed-form transforms, light
radiance and camera field of view, and round trips through the writer.
"""
from __future__ import annotations

import math
import struct

import numpy as np

F = np.float32
PI = F(3.1415926)  # math/utils.h:28


def sid(name: str) -> int:
    """core/strid.h:39-60 (the exporter's strid.py restates the same): CRC-32, reflected
    polynomial 0xEDB88320, initial 0, final complement."""
    crc = 0
    for ch in name.encode("latin-1"):
        crc ^= ch
        for _ in range(8):
            crc = (crc >> 1) ^ (-(crc & 1) & 0xEDB88320)
    return (~crc) & 0xFFFFFFFF


# ---- writer (exporter byte layout) ------------------------------------------------------------

class SortFileWriter:
    """stream.py FileStream.serialize: floats 'f', ints 'I', bools '?', strings ASCII + NUL."""

    def __init__(self):
        self.buf = bytearray()

    def u32(self, v):
        self.buf += struct.pack("<I", v & 0xFFFFFFFF)

    def i32(self, v):
        self.buf += struct.pack("<i", v)

    def f32(self, *vs):
        for v in vs:
            self.buf += struct.pack("<f", v)

    def b8(self, v):
        self.buf += struct.pack("<?", bool(v))

    def s(self, text):
        self.buf += text.encode("ascii") + b"\0"

    def sid(self, name):
        self.u32(sid(name))

    def raw(self, b):
        self.buf += bytes(b)


DIFFUSE, GGX_NODE, GLASS, DISNEY_NODE, OTHER = ("SORTNode_Material_DiffuseLambert", "SORTNode_Material_MicrofacetReflection",
                                                "SORTNode_Material_Glass", "SORTNode_Material_DisneyBRDF",
                                                "SORTNode_Material_Cloth")
TRANSPARENT_NODE = "SORTNode_Material_Transparent"
_SOURCES = {  # TSL bodies stand-ins: only the closure names the loader inspects matter
    DIFFUSE: "shader bxdf_lambert(color Diffuse, vector Normal, out closure Result){ Result = make_closure<lambert>( Diffuse , Normal ); }",
    GGX_NODE: "shader bxdf_ggx(...){ Result = make_closure<microfacet_reflection_ggx>( InteriorIOR , AbsorptionCoefficient , RoughnessU , RoughnessV , BaseColor , Normal ); }",
    GLASS: "shader bxdf_glass(...){ Result = make_closure<dieletric>( Reflectance , Transmittance , RoughnessU , RoughnessV , Normal ); }",
    DISNEY_NODE: "shader bxdf_disney(...){ Result = make_closure<disney>( Metallic , Specular , SpecularTint , Roughness , Anisotropic , Sheen , SheenTint , Clearcoat , ClearcoatGlossiness , SpecularTransmittance , ScatterDistance , Flatness , DiffuseTransmittance , false , BaseColor , Normal ); }",
    OTHER: "shader bxdf_cloth(...){ }",
    TRANSPARENT_NODE: "shader bxdf_transparent(color Attenuation, out closure Result){ Result = make_closure<transparent>( Attenuation ); }",
    "SORTNodeOutput": "shader output(closure Surface, closure Volume){ }",
}


def _node_params(kind, p):
    """(name, value) per serialized input, in the node's serialize_prop order (material.py)."""
    if kind == DIFFUSE:
        return [("Diffuse", p.get("Diffuse", (0.8, 0.8, 0.8))), ("Normal", (0.0, 0.0, 1.0))]
    if kind == GGX_NODE:
        return [("InteriorIOR", p.get("InteriorIOR", (0.37, 0.37, 0.37))),
                ("AbsorptionCoefficient", p.get("AbsorptionCoefficient", (2.82, 2.82, 2.82))),
                ("RoughnessU", p.get("RoughnessU", 0.2)), ("RoughnessV", p.get("RoughnessV", p.get("RoughnessU", 0.2))),
                ("BaseColor", p.get("BaseColor", (1.0, 1.0, 1.0))), ("Normal", (0.0, 0.0, 1.0))]
    if kind == GLASS:
        return [("Reflectance", p.get("Reflectance", (1.0, 1.0, 1.0))),
                ("Transmittance", p.get("Transmittance", (1.0, 1.0, 1.0))),
                ("RoughnessU", p.get("RoughnessU", 0.1)), ("RoughnessV", p.get("RoughnessV", p.get("RoughnessU", 0.1))),
                ("Normal", (0.0, 0.0, 1.0))]
    if kind == DISNEY_NODE:
        d = dict(Metallic=1.0, Specular=1.0, SpecularTint=0.0, Roughness=0.2, Anisotropic=0.0, Sheen=1.0,
                 SheenTint=0.0, Clearcoat=0.0, ClearcoatGlossiness=1.0, SpecularTransmittance=0.0,
                 ScatterDistance=(0.0, 0.0, 0.0), Flatness=0.0, DiffuseTransmittance=0.0, BaseColor=(0.8, 0.8, 0.8))
        d.update(p)
        order = ["Metallic", "Specular", "SpecularTint", "Roughness", "Anisotropic", "Sheen", "SheenTint", "Clearcoat",
                 "ClearcoatGlossiness", "SpecularTransmittance", "ScatterDistance", "Flatness",
                 "DiffuseTransmittance", "BaseColor"]
        return [(k, d[k]) for k in order] + [("Normal", (0.0, 0.0, 1.0))]
    if kind == TRANSPARENT_NODE:
        return [("Attenuation", p.get("Attenuation", (1.0, 1.0, 1.0)))]
    return [("Scale", p.get("Scale", 1.0))]


def _write_params(w, params):
    w.u32(len(params))
    for name, v in params:
        w.s(name)
        if isinstance(v, str):
            w.i32(4)
            w.s(v)
        elif np.ndim(v) == 0:
            w.i32(1)
            w.f32(v)
        else:
            w.i32(3)
            w.f32(*v)


def write_scene(config=None, materials=(), camera=None, meshes=(), lights=(), sky=None, accel=("Obvh", 28, 16),
                resources=(), extra_entities=None) -> bytes:
    """One .sort stream. materials: dicts {name, kind (node type id or None = no surface shader),
    params, transparent, sss, texture_input}; meshes: dicts {matrix (16), verts (n,8) float32
    [position, normal, uv], tris (m,4) int32 [i0, i1, i2, material id], has_uv, volume (w,h,d) or
    None}; lights: dicts {type: AREA/POINT/SPOT/SUN, matrix, energy, color, shape, size, size_y,
    spot (start, range)}; sky: {intensity, color, image}; camera: {eye, up, target, width, height,
    lens, sensor, fit, pixel_aspect, fov}."""
    cfg = dict(threads=8, spp=16, width=64, height=64, clamping=0.0, integrator="PathTracing", max_depth=5,
               bssrdf=2)
    cfg.update(config or {})
    w = SortFileWriter()
    # export_global_config (exporter.py:424-447)
    w.u32(0)
    w.s("/tmp/sort/")
    w.u32(cfg["threads"])
    w.u32(cfg["spp"])
    w.u32(cfg["width"])
    w.u32(cfg["height"])
    w.f32(cfg["clamping"])
    w.sid(cfg["integrator"])
    w.u32(cfg["max_depth"])
    if cfg["integrator"] == "PathTracing":
        w.u32(cfg["bssrdf"])
    elif cfg["integrator"] == "AmbientOcclusion":
        w.f32(cfg.get("ao_max_dist", 1.0))
    elif cfg["integrator"] == "BidirPathTracing":
        w.b8(cfg.get("bdpt_mis", True))
    elif cfg["integrator"] == "InstantRadiosity":
        w.u32(cfg.get("ir_sets", 1))
        w.u32(cfg.get("ir_paths", 64))
        w.f32(cfg.get("ir_min_dist", 1.0))
    # collect_shader_resources (exporter.py:674-723): the renderer reads (file, type) pairs
    w.u32(len(resources))
    for fname, rtype in resources:
        w.s(fname)
        w.sid(rtype)
    # export_materials (exporter.py:726-935)
    seen = set()
    for k, m in enumerate(materials):
        kind = m.get("kind", DIFFUSE)
        if kind is None:
            w.sid("Material")
            w.s(m["name"])
            w.sid("Invalid Surface Shader")
            w.sid("Invalid Volume Shader")
            w.b8(m.get("transparent", False))
            w.b8(m.get("sss", False))
            w.f32(0.1)
            w.u32(1024)
            continue
        bxdf_name = f"{m['name']}_bxdf{k}"
        tex_name = f"{m['name']}_tex{k}"
        nodes = [("ShaderOutput_", "SORTNodeOutput", [])]
        conns = []
        if m.get("texture_input"):
            # a checker texture node feeding the BXDF's first input
            nodes.append((bxdf_name, kind, _node_params(kind, m.get("params", {}))))
            nodes.append((tex_name, "SORTNode_Texture_Checker", [("Color1", (1.0, 1.0, 1.0)), ("Color2", (0.0, 0.0, 0.0))]))
            conns.append((tex_name, "Result", bxdf_name, _node_params(kind, {})[0][0]))
            conns.append((bxdf_name, "Result", "ShaderOutput_" + m["name"], "Surface"))
            units = ["SORTNode_Texture_Checker", kind, "SORTNodeOutput"]
        else:
            nodes.append((bxdf_name, kind, _node_params(kind, m.get("params", {}))))
            conns.append((bxdf_name, "Result", "ShaderOutput_" + m["name"], "Surface"))
            units = [kind, "SORTNodeOutput"]
        for u in units:  # each shader unit type once, sources before the material (exporter.py:866-872)
            if u in seen:
                continue
            seen.add(u)
            w.sid("ShaderUnitTemplate")
            w.s(u)
            w.s(_SOURCES.get(u, "shader texture(out color Result){ }"))
            w.u32(0)
        w.sid("Material")
        w.s(m["name"])
        w.sid("Surface Shader")
        w.u32(len(nodes))
        for name, typ, params in nodes:
            w.s(name)
            w.s(typ)
            _write_params(w, params)
        w.u32(len(conns))
        for c in conns:
            for x in c:
                w.s(x)
        w.sid("Invalid Volume Shader")
        # the exporter sets the flag when the graph holds a transparent node (exporter.py:753-772,926-927)
        w.b8(m.get("transparent", kind == TRANSPARENT_NODE))
        w.b8(m.get("sss", False))
        w.f32(0.1)
        w.u32(1024)
    w.sid("End of Material")
    # export_scene (exporter.py:227-408)
    w.sid("verification bits")
    if camera is not None:
        c = dict(lens=0.0, sensor=(36.0, 24.0), fit=0, pixel_aspect=(1.0, 1.0))
        c.update(camera)
        w.sid("PerspectiveCameraEntity")
        w.f32(*c["eye"])
        w.f32(*c["up"])
        w.f32(*c["target"])
        w.u32(c["width"])
        w.u32(c["height"])
        w.f32(c["lens"])
        w.f32(*c["sensor"])
        w.u32(c["fit"])
        w.f32(*c["pixel_aspect"])
        w.f32(c["fov"])
    for mesh in meshes:
        w.sid("VisualEntity")
        w.f32(*np.asarray(mesh.get("matrix", np.eye(4)), np.float32).reshape(16))
        w.u32(1)
        w.sid("MeshVisual")
        w.b8(mesh.get("has_uv", False))
        verts = np.ascontiguousarray(mesh["verts"], np.float32)
        tris = np.ascontiguousarray(mesh["tris"], np.int32)
        w.i32(len(verts))
        w.raw(verts.tobytes())
        w.i32(len(tris))
        w.raw(tris.tobytes())
        vol = mesh.get("volume")
        if vol is None:
            w.sid("no_volume")
        else:
            w.sid("has_volume")
            w.u32(vol[0])
            w.u32(vol[1])
            w.u32(vol[2])
            w.raw(np.zeros(vol[0] * vol[1] * vol[2], np.float32).tobytes())
        w.sid("end of mesh")
    for hair in (extra_entities or {}).get("hair", []):
        w.sid("VisualEntity")
        w.f32(*np.eye(4, dtype=np.float32).reshape(16))
        w.u32(1)
        w.sid("HairVisual")
        w.u32(len(hair))
        w.f32(0.01, 0.02)
        w.i32(-1)
        for strand in hair:
            w.u32(len(strand) - 1)
            for p in strand:
                w.f32(*p)
    mapping = {"SUN": "DirLightEntity", "POINT": "PointLightEntity", "SPOT": "SpotLightEntity", "AREA": "AreaLightEntity"}
    for lt in lights:
        w.sid(mapping[lt["type"]])
        w.f32(*np.asarray(lt["matrix"], np.float32).reshape(16))
        w.f32(lt["energy"])
        w.f32(*lt["color"])
        if lt["type"] == "SPOT":
            w.f32(*lt.get("spot", (30.0, 45.0)))
        elif lt["type"] == "AREA":
            w.sid(lt["shape"])
            if lt["shape"] == "SQUARE":
                w.f32(lt["size"])
            elif lt["shape"] == "RECTANGLE":
                w.f32(lt["size"])
                w.f32(lt["size_y"])
            elif lt["shape"] == "DISK":
                w.f32(lt["size"] * 0.5)
    if sky is not None:
        w.sid("SkyLightEntity")
        w.f32(*np.eye(4, dtype=np.float32).reshape(16))
        w.f32(sky["intensity"])
        w.f32(*sky["color"])
        w.s(sky.get("image", ""))
    w.sid("End of Entities")
    name, depth, leaf = accel
    w.sid(name)
    if name in ("Bvh", "KDTree", "OcTree", "Qbvh", "Obvh"):
        w.u32(depth)
        w.u32(leaf)
    return bytes(w.buf)


# ---- reader (the renderer's reading, flattened) ----------------------------------------------

def determinant(m):
    """Matrix::Determinant (math/matrix.cpp:104-112), float32, left to right."""
    a = (m[0] * m[5] - m[1] * m[4]) * (m[10] * m[15] - m[11] * m[14])
    b = (m[0] * m[6] - m[2] * m[4]) * (m[9] * m[15] - m[11] * m[13])
    c = (m[0] * m[7] - m[3] * m[4]) * (m[9] * m[14] - m[10] * m[13])
    d = (m[1] * m[6] - m[2] * m[5]) * (m[8] * m[15] - m[11] * m[12])
    e = (m[1] * m[7] - m[3] * m[5]) * (m[8] * m[14] - m[10] * m[12])
    f = (m[2] * m[7] - m[3] * m[6]) * (m[8] * m[13] - m[9] * m[12])
    return a - b + c + d - e + f


def inverse(m):
    """Matrix::Inverse (math/matrix.cpp:115-143), float32; None when |det| < 1e-8."""
    m = [F(x) for x in m]
    d = determinant(m)
    if abs(d) < F(0.00000001):
        return None
    d = F(1.0) / d
    o = [None] * 16
    o[0] = d * (m[5] * (m[10] * m[15] - m[11] * m[14]) + m[6] * (m[11] * m[13] - m[9] * m[15]) + m[7] * (m[9] * m[14] - m[10] * m[13]))
    o[1] = d * (m[9] * (m[2] * m[15] - m[3] * m[14]) + m[10] * (m[3] * m[13] - m[1] * m[15]) + m[11] * (m[1] * m[14] - m[2] * m[13]))
    o[2] = d * (m[13] * (m[2] * m[7] - m[3] * m[6]) + m[14] * (m[3] * m[5] - m[1] * m[7]) + m[15] * (m[1] * m[6] - m[2] * m[5]))
    o[3] = d * (m[1] * (m[7] * m[10] - m[6] * m[11]) + m[2] * (m[5] * m[11] - m[7] * m[9]) + m[3] * (m[6] * m[9] - m[5] * m[10]))
    o[4] = d * (m[6] * (m[8] * m[15] - m[11] * m[12]) + m[7] * (m[10] * m[12] - m[8] * m[14]) + m[4] * (m[11] * m[14] - m[10] * m[15]))
    o[5] = d * (m[10] * (m[0] * m[15] - m[3] * m[12]) + m[11] * (m[2] * m[12] - m[0] * m[14]) + m[8] * (m[3] * m[14] - m[2] * m[15]))
    o[6] = d * (m[14] * (m[0] * m[7] - m[3] * m[4]) + m[15] * (m[2] * m[4] - m[0] * m[6]) + m[12] * (m[3] * m[6] - m[2] * m[7]))
    o[7] = d * (m[2] * (m[7] * m[8] - m[4] * m[11]) + m[3] * (m[4] * m[10] - m[6] * m[8]) + m[0] * (m[6] * m[11] - m[7] * m[10]))
    o[8] = d * (m[7] * (m[8] * m[13] - m[9] * m[12]) + m[4] * (m[9] * m[15] - m[11] * m[13]) + m[5] * (m[11] * m[12] - m[8] * m[15]))
    o[9] = d * (m[11] * (m[0] * m[13] - m[1] * m[12]) + m[8] * (m[1] * m[15] - m[3] * m[13]) + m[9] * (m[3] * m[12] - m[0] * m[15]))
    o[10] = d * (m[15] * (m[0] * m[5] - m[1] * m[4]) + m[12] * (m[1] * m[7] - m[3] * m[5]) + m[13] * (m[3] * m[4] - m[0] * m[7]))
    o[11] = d * (m[3] * (m[5] * m[8] - m[4] * m[9]) + m[0] * (m[7] * m[9] - m[5] * m[11]) + m[1] * (m[4] * m[11] - m[7] * m[8]))
    o[12] = d * (m[4] * (m[10] * m[13] - m[9] * m[14]) + m[5] * (m[8] * m[14] - m[10] * m[12]) + m[6] * (m[9] * m[12] - m[8] * m[13]))
    o[13] = d * (m[8] * (m[2] * m[13] - m[1] * m[14]) + m[9] * (m[0] * m[14] - m[2] * m[12]) + m[10] * (m[1] * m[12] - m[0] * m[13]))
    o[14] = d * (m[12] * (m[2] * m[5] - m[1] * m[6]) + m[13] * (m[0] * m[6] - m[2] * m[4]) + m[14] * (m[1] * m[4] - m[0] * m[5]))
    o[15] = d * (m[0] * (m[5] * m[10] - m[6] * m[9]) + m[1] * (m[6] * m[8] - m[4] * m[10]) + m[2] * (m[4] * m[9] - m[5] * m[8]))
    return np.array(o, np.float32)


def matmul(a, b):
    """Matrix::operator* (math/matrix.cpp:51-67): sums from 0.0f over k."""
    r = np.zeros(16, np.float32)
    for i in range(4):
        for j in range(4):
            acc = F(0.0)
            for k in range(4):
                acc = acc + F(a[i * 4 + k]) * F(b[k * 4 + j])
            r[i * 4 + j] = acc
    return r


class _R:
    def __init__(self, data):
        self.d = memoryview(data)
        self.o = 0

    def take(self, n):
        if self.o + n > len(self.d):
            raise ValueError(f"truncated at byte {self.o}")
        b = self.d[self.o:self.o + n]
        self.o += n
        return b

    def u32(self):
        return struct.unpack("<I", self.take(4))[0]

    def i32(self):
        return struct.unpack("<i", self.take(4))[0]

    def f32(self, n=1):
        v = np.frombuffer(self.take(4 * n), "<f4").astype(np.float32)
        return v[0] if n == 1 else v

    def b8(self):
        return self.take(1)[0] != 0

    def s(self):
        end = bytes(self.d[self.o:]).find(b"\0")
        if end < 0:
            raise ValueError(f"unterminated string at byte {self.o}")
        t = bytes(self.d[self.o:self.o + end]).decode("latin-1")
        self.o += end + 1
        return t

    def transform(self):
        m = self.f32(16)
        inv = inverse(m)
        return m, (inv if inv is not None else np.eye(4, dtype=np.float32).reshape(16))


def _graph(r):
    nodes, conns = [], []
    for _ in range(r.u32()):
        name, typ = r.s(), r.s()
        params = {}
        for _ in range(r.u32()):
            pname, ch = r.s(), r.i32()
            if ch == 1:
                params[pname] = (1, [r.f32()])
            elif ch == 3:
                params[pname] = (3, list(r.f32(3)))
            elif ch == 4:
                params[pname] = (4, r.s())
        nodes.append((name, typ, params))
    for _ in range(r.u32()):
        conns.append((r.s(), r.s(), r.s(), r.s()))
    return nodes, conns


LAMBERT, GGX, DIELECTRIC, DISNEY, TRANSPARENT = 0, 1, 2, 3, 4


def _white():
    return dict(type=LAMBERT, base=(1.0, 1.0, 1.0), trans=(0.0, 0.0, 0.0), roughness=0.0, ior=0.0, eta=(0.0, 0.0, 0.0),
                k=(0.0, 0.0, 0.0), metallic=0.0, specular=0.0, specular_tint=0.0, sheen=0.0, sheen_tint=0.0,
                clearcoat=0.0, clearcoat_gloss=0.0)


def _map(mat, sources):
    """(GPU material dict, 'mapped' | 'approx' | 'unmapped') for one material's surface graph."""
    out = _white()
    if mat["surface"] is None:
        return out, "mapped"
    nodes, conns = mat["surface"]
    outn = [n for n in nodes if n[1] == "SORTNodeOutput"]
    if not outn:
        return out, "unmapped"
    oname = outn[-1][0]
    bxdf = None
    for c in conns:
        if c[2] in (oname, oname + mat["name"]):
            for n in nodes:
                if n[0] == c[0]:
                    bxdf = n
    if bxdf is None:
        return out, "unmapped"
    approx = (mat["transparent"] and bxdf[1] != "SORTNode_Material_Transparent") or mat["sss"]
    if any(c[2] == bxdf[0] and c[3] != "Normal" for c in conns):
        approx = True
    params = bxdf[2]

    def f1(name, d):
        p = params.get(name)
        return float(p[1][0]) if p and p[0] == 1 else d

    def f3(name, d):
        p = params.get(name)
        return tuple(float(x) for x in p[1]) if p and p[0] == 3 else (d, d, d)

    t = bxdf[1]
    src = sources.get(t)
    if t == "SORTNode_Material_DiffuseLambert":
        out.update(type=LAMBERT, base=f3("Diffuse", 1.0))
    elif t.startswith("SORTNode_Material_MicrofacetReflection"):
        out.update(type=GGX, base=f3("BaseColor", 1.0), eta=f3("InteriorIOR", 0.37), k=f3("AbsorptionCoefficient", 2.82),
                   roughness=f1("RoughnessU", 0.0))
        if f1("RoughnessV", out["roughness"]) != out["roughness"] or src is None or "microfacet_reflection_ggx" not in src:
            approx = True
    elif t == "SORTNode_Material_Glass":
        out.update(type=DIELECTRIC, base=f3("Reflectance", 1.0), trans=f3("Transmittance", 1.0),
                   roughness=f1("RoughnessU", 0.0), ior=1.5)
        if f1("RoughnessV", out["roughness"]) != out["roughness"]:
            approx = True
    elif t.startswith("SORTNode_Material_DisneyBRDF"):
        out.update(type=DISNEY, base=f3("BaseColor", 1.0), metallic=f1("Metallic", 1.0), specular=f1("Specular", 1.0),
                   specular_tint=f1("SpecularTint", 0.0), roughness=f1("Roughness", 0.2), sheen=f1("Sheen", 1.0),
                   sheen_tint=f1("SheenTint", 0.0), clearcoat=f1("Clearcoat", 0.0),
                   clearcoat_gloss=f1("ClearcoatGlossiness", 1.0))
        sd = f3("ScatterDistance", 0.0)
        if (f1("Anisotropic", 0.0) != 0 or f1("SpecularTransmittance", 0.0) != 0 or f1("DiffuseTransmittance", 0.0) != 0
                or max(sd) > 0 or src is None or ", true ," in src):
            approx = True
    elif t == "SORTNode_Material_Transparent":
        out.update(type=TRANSPARENT, base=f3("Attenuation", 1.0))
    else:
        return out, "unmapped"
    return out, ("approx" if approx else "mapped")


def _normalize_rows(v):
    """Vector::Normalize (math/vector3.h:104-109) of each row: divided by its float32 length
    unless that is zero."""
    v = v.astype(np.float32)
    ln = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
    nz = ln != F(0.0)
    v[nz] = v[nz] / ln[nz, None]
    return v


def gen_uv(p):
    """Mesh::GenUV (core/mesh.cpp:60-74) for a mesh stored without texture coordinates: for each
    world-space vertex, the normalised direction d from the mean vertex (a float32 running sum
    divided by the count) mapped to (acos(clamp(d.y)) * INV_PI, phi * INV_TWOPI), phi = atan2(d.z,
    d.x) lifted into [0, 2pi) (sphericalTheta / sphericalPhi, bxdf_utils.h:80-87; constants of
    math/utils.h:27-33). acos and atan2 are taken in double and rounded to float32."""
    n = len(p)
    if n == 0:
        return np.zeros((0, 2), np.float32)
    c = np.cumsum(p.astype(np.float32), axis=0, dtype=np.float32)[-1]  # sequential float32 sum
    c = (c / F(n)).astype(np.float32)
    d = _normalize_rows(p - c)
    y = np.clip(d[:, 1], F(-1.0), F(1.0))
    theta = np.arccos(y.astype(np.float64)).astype(np.float32)
    phi = np.arctan2(d[:, 2].astype(np.float64), d[:, 0].astype(np.float64)).astype(np.float32)
    phi = np.where(phi < F(0.0), phi + F(2.0) * PI, phi).astype(np.float32)
    return np.stack([theta * F(0.3183099), phi * F(0.15915494)], 1).astype(np.float32)


def tri_tangents(p, uv, tris):
    """Mesh::genTagentForTri (core/mesh.cpp:76-108) for every triangle (rows of vertex indices):
    (dv2 dp1 - dv1 dp2) / det with du1 = u0 - u2, du2 = u1 - u2, dv1 = v0 - v2, dv2 = v1 - v2,
    dp1 = p0 - p2, dp2 = p1 - p2, det = du1 dv2 - dv1 du2; for det == 0 the first axis of
    coordinateSystem (math/vector3.h:171-180) about normalize(cross(p0 - p1, p0 - p2))."""
    p0, p1, p2 = p[tris[:, 0]], p[tris[:, 1]], p[tris[:, 2]]
    t0, t1, t2 = uv[tris[:, 0]], uv[tris[:, 1]], uv[tris[:, 2]]
    du1, du2 = t0[:, 0] - t2[:, 0], t1[:, 0] - t2[:, 0]
    dv1, dv2 = t0[:, 1] - t2[:, 1], t1[:, 1] - t2[:, 1]
    det = du1 * dv2 - dv1 * du2
    dp1, dp2 = p0 - p2, p1 - p2
    with np.errstate(divide="ignore", invalid="ignore"):
        out = (dv2[:, None] * dp1 - dv1[:, None] * dp2) / det[:, None]
        a, b = p0 - p1, p0 - p2
        n = np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1], a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                      a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], 1).astype(np.float32)
        n = _normalize_rows(n)
        xz = F(1.0) / np.sqrt(n[:, 0] * n[:, 0] + n[:, 2] * n[:, 2])
        yz = F(1.0) / np.sqrt(n[:, 1] * n[:, 1] + n[:, 2] * n[:, 2])
        zero = np.zeros(len(n), np.float32)
        ax = np.stack([-n[:, 2] * xz, zero, n[:, 0] * xz], 1)
        ay = np.stack([zero, n[:, 2] * yz, -n[:, 1] * yz], 1)
        frame = np.where((np.abs(n[:, 0]) > np.abs(n[:, 1]))[:, None], ax, ay)
    return np.where((det == F(0.0))[:, None], frame, out).astype(np.float32)


def smooth_tangents(p, uv, tris):
    """Mesh::GenSmoothTagent (core/mesh.cpp:42-58): per vertex, the float32 sum in triangle order
    of the tangents of the triangles that name it (once per corner), normalised."""
    acc = np.zeros((len(p), 3), np.float32)
    if len(tris):
        t = tri_tangents(p, uv, tris.astype(np.int64))
        flat = tris.astype(np.int64).reshape(-1)
        # unbuffered: a vertex named again is added again, in index order (triangle order)
        np.add.at(acc, flat, np.repeat(t, 3, axis=0))
    return _normalize_rows(acc)


def read_scene(data: bytes) -> dict:
    """The scene dictionary the C ABI loader returns for the same bytes (see module docstring)."""
    r = _R(data)
    skipped = {}

    def flag(name, n=1):
        skipped[name] = skipped.get(name, 0) + n

    if r.u32() != 0:
        raise ValueError("unsupported configuration version")
    r.s()
    cfg = dict(threads=r.u32(), spp=r.u32(), width=r.u32(), height=r.u32(), clamping=float(r.f32()), integrator=r.u32(),
               max_recursive_depth=0, max_bssrdf_bounces=0)
    known = {sid(n): n for n in ("PathTracing", "AmbientOcclusion", "BidirPathTracing", "LightTracing",
                                 "InstantRadiosity", "DirectLight", "WhittedRT")}
    it = known.get(cfg["integrator"])
    if it is not None:
        cfg["max_recursive_depth"] = r.i32()
        if it == "PathTracing":
            cfg["max_bssrdf_bounces"] = r.i32()
        elif it == "AmbientOcclusion":
            r.f32()
        elif it == "BidirPathTracing":
            r.b8()
        elif it == "InstantRadiosity":
            r.i32(), r.i32(), r.f32()
    for _ in range(r.u32()):
        r.s(), r.u32()
    sources, mats = {}, []
    while True:
        tag = r.u32()
        if tag == sid("End of Material"):
            break
        if tag == sid("ShaderUnitTemplate"):
            typ, src = r.s(), r.s()
            sources.setdefault(typ, src)
            for _ in range(r.u32()):
                r.s(), r.s()
        elif tag == sid("ShaderGroupTemplate"):
            r.s()
            _graph(r)
            r.s()
            for _ in range(r.u32()):
                r.s()
            if r.s():
                for _ in range(r.u32()):
                    r.s()
        elif tag == sid("Material"):
            m = dict(name=r.s(), surface=None, volume=False)
            if r.u32() == sid("Surface Shader"):
                m["surface"] = _graph(r)
            if r.u32() == sid("Volume Shader"):
                _graph(r)
                m["volume"] = True
            m["transparent"], m["sss"] = r.b8(), r.b8()
            r.f32(), r.u32()
            mats.append(m)
        else:
            raise ValueError(f"unknown material record at byte {r.o - 4}")
    materials, names = [], []
    for m in mats:
        gm, how = _map(m, sources)
        if how == "unmapped":
            flag("unmapped_materials")
        elif how == "approx":
            flag("approximated_materials")
        if m["volume"]:
            flag("volumes")
        materials.append(gm)
        names.append(m["name"])
    materials.append(_white())
    names.append("default")
    nm = len(mats)

    if r.u32() != sid("verification bits"):
        raise ValueError("not a scene stream")
    pos, nrm, uvs, tans, idx, tmat = [], [], [], [], [], []
    quads, qrad = [], []
    env = np.zeros(3, np.float32)
    camera = None
    any_light = False
    nv_total = 0
    while True:
        cls = r.u32()
        if cls == sid("End of Entities"):
            break
        if cls == sid("PerspectiveCameraEntity"):
            camera = dict(eye=r.f32(3), up=r.f32(3), target=r.f32(3), width=r.u32(), height=r.u32(), lens=r.f32(),
                          sensor=(r.f32(), r.f32()), fit=r.i32(), pixel_aspect=(r.f32(), r.f32()), fov=r.f32())
        elif cls == sid("VisualEntity"):
            m, inv = r.transform()
            for _ in range(r.u32()):
                vc = r.u32()
                if vc == sid("MeshVisual"):
                    has_uv = r.b8()
                    nv = r.u32()
                    v = np.frombuffer(bytes(r.take(32 * nv)), "<f4").reshape(nv, 8).astype(np.float32)
                    p, n, t = v[:, 0:3], v[:, 3:6].copy(), v[:, 6:8]
                    # Matrix::TransformPoint (math/matrix.cpp:70-82)
                    x = p[:, 0] * m[0] + p[:, 1] * m[1] + p[:, 2] * m[2] + m[3]
                    y = p[:, 0] * m[4] + p[:, 1] * m[5] + p[:, 2] * m[6] + m[7]
                    z = p[:, 0] * m[8] + p[:, 1] * m[9] + p[:, 2] * m[10] + m[11]
                    w_ = p[:, 0] * m[12] + p[:, 1] * m[13] + p[:, 2] * m[14] + m[15]
                    sel = w_ != F(1.0)
                    x = np.where(sel, x / w_, x)
                    y = np.where(sel, y / w_, y)
                    z = np.where(sel, z / w_, z)
                    pos.append(np.stack([x, y, z], 1).astype(np.float32))
                    # Vector::Normalize then Transform::TransformNormal (transposed inverse)
                    ln = np.sqrt(n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1] + n[:, 2] * n[:, 2])
                    nz = ln != F(0.0)
                    n[nz] = n[nz] / ln[nz, None]
                    nx = n[:, 0] * inv[0] + n[:, 1] * inv[4] + n[:, 2] * inv[8]
                    ny = n[:, 0] * inv[1] + n[:, 1] * inv[5] + n[:, 2] * inv[9]
                    nzz = n[:, 0] * inv[2] + n[:, 1] * inv[6] + n[:, 2] * inv[10]
                    nrm.append(np.stack([nx, ny, nzz], 1).astype(np.float32))
                    nt = r.u32()
                    tr = np.frombuffer(bytes(r.take(16 * nt)), "<i4").reshape(nt, 4)
                    if nt and (tr[:, :3].astype(np.int64).min() < 0 or tr[:, :3].astype(np.int64).max() >= nv):
                        raise ValueError("vertex index out of range")
                    t = t.copy() if has_uv else gen_uv(pos[-1])
                    uvs.append(t)
                    tans.append(smooth_tangents(pos[-1], t, tr[:, :3]))
                    idx.append((tr[:, :3].astype(np.uint32) + np.uint32(nv_total)))
                    mi = tr[:, 3]
                    tmat.append(np.where((mi >= 0) & (mi < nm), mi, nm).astype(np.uint32))
                    nv_total += nv
                    vol = r.u32()
                    if vol == sid("has_volume"):
                        a, b, c = r.u32(), r.u32(), r.u32()
                        if a and b and c:
                            r.take(4 * a * b * c)
                        flag("volumes")
                    elif vol != sid("no_volume"):
                        raise ValueError("unknown volume tag")
                    if r.u32() != sid("end of mesh"):
                        raise ValueError("missing end of mesh")
                elif vc == sid("HairVisual"):
                    n_h = r.u32()
                    r.f32(), r.f32(), r.i32()
                    for _ in range(n_h):
                        st = r.u32()
                        if st:
                            r.take(12 * (st + 1))
                    flag("hair", n_h)
                else:
                    raise ValueError("unknown visual")
        elif cls in (sid("PointLightEntity"), sid("DirLightEntity"), sid("SpotLightEntity")):
            r.transform()
            r.f32(), r.f32(3)
            if cls == sid("SpotLightEntity"):
                r.f32(), r.f32()
            any_light = True
            flag("punctual_lights")
        elif cls == sid("SkyLightEntity"):
            r.transform()
            inten = r.f32()
            col = r.f32(3)
            img = r.s()
            env = (col * inten).astype(np.float32)
            if img:
                flag("sky_image")
            any_light = True
        elif cls == sid("AreaLightEntity"):
            # AreaLightEntity::Serialize (light_entity.cpp:99-155)
            m, inv = r.transform()
            sx = np.sqrt(m[0] * m[0] + m[4] * m[4] + m[8] * m[8])
            sy = np.sqrt(m[1] * m[1] + m[5] * m[5] + m[9] * m[9])
            sz = np.sqrt(m[2] * m[2] + m[6] * m[6] + m[10] * m[10])
            down = np.eye(4, dtype=np.float32).reshape(16)
            down[0], down[5], down[10] = F(1) / sx, F(1) / sy, F(1) / sz
            up = np.eye(4, dtype=np.float32).reshape(16)
            up[0], up[5], up[10] = sx, sy, sz
            mat, minv = matmul(m, down), matmul(up, inv)
            energy = r.f32()
            col = r.f32(3)
            shape = r.u32()
            any_light = True
            if shape == sid("SQUARE"):
                size = r.f32()
                sxy = (size, size)
                scale = F(0.8) * energy / (size * size * sx * sy * PI)
            elif shape == sid("RECTANGLE"):
                a, b = r.f32(), r.f32()
                sxy = (a, b)
                scale = F(0.8) * energy / (a * sx * b * sy * PI)
            elif shape == sid("DISK"):
                r.f32()
                flag("disk_lights")
                continue
            else:
                continue
            quads.append({"mat": mat.reshape(4, 4), "inv": minv.reshape(4, 4), "size_x": float(sxy[0]),
                          "size_y": float(sxy[1])})
            qrad.append((col * scale).astype(np.float32))
        else:
            raise ValueError(f"unknown entity at byte {r.o - 4}")
    if not any_light:
        env = np.full(3, F(0.5), np.float32)
    acc = r.u32()
    depth = leaf = 0
    if acc in {sid(n) for n in ("Bvh", "KDTree", "OcTree", "Qbvh", "Obvh")}:
        depth, leaf = r.i32(), r.i32()
    cfg.update(accelerator=acc, accel_max_depth=depth, accel_max_leaf=leaf)
    cam = None
    if camera is not None:
        c = camera
        w, h = F(c["width"]), F(c["height"])
        aspect = w / h * c["pixel_aspect"][0] / c["pixel_aspect"][1]
        ys = F(1.0) / F(math.tan(F(c["fov"]) * F(0.5)))
        if c["sensor"][0] != 0 and c["sensor"][1] != 0 and (c["fit"] == 1 or (c["fit"] != 2 and aspect > 1)):
            ys = ys * aspect
        fov_y = F(2.0) * F(math.atan(F(1.0) / ys))
        cam = dict(eye=tuple(float(x) for x in c["eye"]), target=tuple(float(x) for x in c["target"]),
                   up=tuple(float(x) for x in c["up"]), fov_y_deg=math.degrees(float(fov_y)))
        if c["lens"] != 0 or c["pixel_aspect"][0] != c["pixel_aspect"][1]:
            flag("camera_lens_or_pixel_aspect")

    def cat(parts, shape, dt):
        return np.concatenate(parts).astype(dt) if parts else np.zeros(shape, dt)

    idx_a = cat(idx, (0, 3), np.uint32)
    return {"pos": cat(pos, (0, 3), np.float32), "idx": idx_a, "quads": quads,
            "normals": cat(nrm, (0, 3), np.float32), "uv": cat(uvs, (0, 2), np.float32),
            "tangents": cat(tans, (0, 3), np.float32),
            "tri_material": cat(tmat, (0,), np.uint32), "materials": materials, "material_names": names,
            "quad_radiance": np.array(qrad, np.float32).reshape(-1, 3), "env": tuple(float(x) for x in env),
            "ntris": int(len(idx_a)), "camera": cam, "config": cfg, "skipped": skipped}


def export_scene_dict(sc, width=64, height=64, spp=16, max_depth=5, energy_scale=None) -> bytes:
    """A .sort stream for one of the renderer's scene dictionaries (``scenes.cornell_sphere()``
    form): one mesh (identity transform, per-vertex normals), one Diffuse / MicrofacetReflection /
    Glass / Disney node per material, one square area light per quad (its transform unscaled, the
    energy chosen so that the loaded radiance is the dictionary's up to float rounding), the
    constant environment as a sky, the camera with its vertical field of view."""
    kinds = {LAMBERT: DIFFUSE, GGX: GGX_NODE, DIELECTRIC: GLASS, DISNEY: DISNEY_NODE}
    mats = []
    for k, m in enumerate(sc["materials"]):
        t = m["type"]
        if t == LAMBERT:
            p = {"Diffuse": tuple(m["base"])}
        elif t == GGX:
            p = {"BaseColor": tuple(m["base"]), "InteriorIOR": tuple(m["eta"]), "AbsorptionCoefficient": tuple(m["k"]),
                 "RoughnessU": m["roughness"]}
        elif t == DIELECTRIC:
            p = {"Reflectance": tuple(m["base"]), "Transmittance": tuple(m["trans"]), "RoughnessU": m["roughness"]}
        else:
            p = {"BaseColor": tuple(m["base"]), "Metallic": m["metallic"], "Specular": m["specular"],
                 "SpecularTint": m["specular_tint"], "Roughness": m["roughness"], "Sheen": m["sheen"],
                 "SheenTint": m["sheen_tint"], "Clearcoat": m["clearcoat"], "ClearcoatGlossiness": m["clearcoat_gloss"]}
        mats.append({"name": f"mat{k}", "kind": kinds[t], "params": p})
    pos = np.asarray(sc["pos"], np.float32)
    nrm = np.asarray(sc["normals"], np.float32)
    verts = np.concatenate([pos, nrm, np.zeros((len(pos), 2), np.float32)], 1)
    tris = np.concatenate([np.asarray(sc["idx"], np.int64), np.asarray(sc["tri_material"], np.int64)[:, None]], 1)
    lights = []
    for q, rad in zip(sc["quads"], sc["quad_radiance"]):
        size = F(q["size_x"])
        e = F(float(size) * float(size) * float(PI) / 0.8)  # 0.8 E / (size^2 pi) ~ 1
        lights.append({"type": "AREA", "matrix": np.asarray(q["mat"], np.float32).reshape(16), "energy": e,
                       "color": tuple(float(x) for x in rad), "shape": "SQUARE", "size": float(size)})
    env = tuple(sc.get("env", (0.0, 0.0, 0.0)))
    sky = {"intensity": 1.0, "color": env} if any(x > 0 for x in env) else None
    cam = sc["camera"]
    camera = {"eye": cam["eye"], "up": cam.get("up", (0.0, 1.0, 0.0)), "target": cam["target"], "width": width,
              "height": height, "sensor": (0.0, 0.0), "fov": math.radians(cam["fov_y_deg"])}
    return write_scene(config={"spp": spp, "width": width, "height": height, "max_depth": max_depth},
                       materials=mats, camera=camera,
                       meshes=[{"verts": verts, "tris": tris, "matrix": np.eye(4, dtype=np.float32)}],
                       lights=lights, sky=sky)
