// Reader of SORT's .sort scene stream (host code): global configuration, material section and
// entity stream, flattened into the arrays that sortgpu_accel_build and sortgpu_scene_create take.
//
// The stream is little-endian binary with no framing: 4-byte unsigned/int/float values, strings
// NUL-terminated, bools 1 byte, a StringID as its 4-byte CRC (core/strid.h:39-60), a Transform
// as 16 floats whose inverse the reader computes (stream/stream.h:180-215), a Spectrum as 3
// floats. The order of the reads is the reference readers' (cited per record below). Every read
// is bounds-checked; where SORT asserts on a broken stream, the loader refuses the file with the
// byte offset.
//
// Arithmetic that produces world-space data (point and normal transforms, the inverse of every
// Transform, the area-light normalisation and radiance) follows the reference's expressions in
// the same operation order, so that positions, normals and quad matrices are the floats SORT
// computes (the library is compiled without FMA contraction).

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "api_internal.h"
#include "sortgpu.h"

namespace {

// StringID: bitwise CRC-32 with polynomial 0xEDB88320, initial value 0 and a final complement
// (core/strid.h:39-60). Characters enter as signed chars, as in the reference.
constexpr uint32_t sid(const char* s) {
    uint32_t crc = 0;
    for (; *s; ++s) {
        crc ^= (uint32_t)(int32_t)(signed char)*s;
        for (int b = 0; b < 8; b++) crc = (crc >> 1) ^ ((0u - (crc & 1u)) & 0xEDB88320u);
    }
    return ~crc;
}

constexpr float kPi = 3.1415926f;  // math/utils.h:28

struct Mat4 {
    float m[16];
};

Mat4 identity() {
    Mat4 r{};
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    return r;
}

// Matrix::operator* (math/matrix.cpp:51-67): each entry accumulated from 0.0f over k.
Mat4 mul(const Mat4& a, const Mat4& b) {
    Mat4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float acc = 0.0f;
            for (int k = 0; k < 4; k++) acc += a.m[i * 4 + k] * b.m[k * 4 + j];
            r.m[i * 4 + j] = acc;
        }
    return r;
}

// Matrix::Determinant (math/matrix.cpp:104-112): six 2x2 minor products, summed left to right.
float determinant(const float* m) {
    const float a = (m[0] * m[5] - m[1] * m[4]) * (m[10] * m[15] - m[11] * m[14]);
    const float b = (m[0] * m[6] - m[2] * m[4]) * (m[9] * m[15] - m[11] * m[13]);
    const float c = (m[0] * m[7] - m[3] * m[4]) * (m[9] * m[14] - m[10] * m[13]);
    const float d = (m[1] * m[6] - m[2] * m[5]) * (m[8] * m[15] - m[11] * m[12]);
    const float e = (m[1] * m[7] - m[3] * m[5]) * (m[8] * m[14] - m[10] * m[12]);
    const float f = (m[2] * m[7] - m[3] * m[6]) * (m[8] * m[13] - m[9] * m[12]);
    return a - b + c + d - e + f;
}

// Matrix::Inverse (math/matrix.cpp:115-143): the adjugate by cofactors over 1/det. Each entry is
// x0*(x1*x2 - x3*x4) + x5*(x6*x7 - x8*x9) + x10*(x11*x12 - x13*x14) with the indices below, in
// the reference's order of terms and factors (the rounding of the result depends on it).
constexpr unsigned char kCofactor[16][15] = {
    {5, 10, 15, 11, 14, 6, 11, 13, 9, 15, 7, 9, 14, 10, 13},
    {9, 2, 15, 3, 14, 10, 3, 13, 1, 15, 11, 1, 14, 2, 13},
    {13, 2, 7, 3, 6, 14, 3, 5, 1, 7, 15, 1, 6, 2, 5},
    {1, 7, 10, 6, 11, 2, 5, 11, 7, 9, 3, 6, 9, 5, 10},
    {6, 8, 15, 11, 12, 7, 10, 12, 8, 14, 4, 11, 14, 10, 15},
    {10, 0, 15, 3, 12, 11, 2, 12, 0, 14, 8, 3, 14, 2, 15},
    {14, 0, 7, 3, 4, 15, 2, 4, 0, 6, 12, 3, 6, 2, 7},
    {2, 7, 8, 4, 11, 3, 4, 10, 6, 8, 0, 6, 11, 7, 10},
    {7, 8, 13, 9, 12, 4, 9, 15, 11, 13, 5, 11, 12, 8, 15},
    {11, 0, 13, 1, 12, 8, 1, 15, 3, 13, 9, 3, 12, 0, 15},
    {15, 0, 5, 1, 4, 12, 1, 7, 3, 5, 13, 3, 4, 0, 7},
    {3, 5, 8, 4, 9, 0, 7, 9, 5, 11, 1, 4, 11, 7, 8},
    {4, 10, 13, 9, 14, 5, 8, 14, 10, 12, 6, 9, 12, 8, 13},
    {8, 2, 13, 1, 14, 9, 0, 14, 2, 12, 10, 1, 12, 0, 13},
    {12, 2, 5, 1, 6, 13, 0, 6, 2, 4, 14, 1, 4, 0, 5},
    {0, 5, 10, 6, 9, 1, 6, 8, 4, 10, 2, 4, 9, 5, 8},
};

// Returns false (out untouched) for |det| < 1e-8, as the reference does.
bool inverse(const Mat4& a, Mat4& out) {
    const float* m = a.m;
    float d = determinant(m);
    if (std::fabs(d) < 0.00000001f) return false;
    d = 1.0f / d;
    for (int e = 0; e < 16; e++) {
        const unsigned char* x = kCofactor[e];
        const float t0 = m[x[0]] * (m[x[1]] * m[x[2]] - m[x[3]] * m[x[4]]);
        const float t1 = m[x[5]] * (m[x[6]] * m[x[7]] - m[x[8]] * m[x[9]]);
        const float t2 = m[x[10]] * (m[x[11]] * m[x[12]] - m[x[13]] * m[x[14]]);
        out.m[e] = d * (t0 + t1 + t2);
    }
    return true;
}

struct Transform {
    Mat4 mat = identity();
    Mat4 inv = identity();
};

// Vector::Length (math/vector3.h:90-99).
float length3(float x, float y, float z) { return std::sqrt(x * x + y * y + z * z); }

// ---- bounds-checked reader --------------------------------------------------------------------

struct FormatError {
    std::string msg;
};

class Reader {
public:
    Reader(const uint8_t* p, size_t n) : p_(p), n_(n) {}
    size_t remaining() const { return n_ - off_; }
    uint32_t u32(const char* what) {
        uint32_t v;
        take(&v, 4, what);
        return v;
    }
    int32_t i32(const char* what) {
        int32_t v;
        take(&v, 4, what);
        return v;
    }
    float f32(const char* what) {
        float v;
        take(&v, 4, what);
        return v;
    }
    bool b8(const char* what) {
        uint8_t v;
        take(&v, 1, what);
        return v != 0;
    }
    std::string str(const char* what) {
        const void* z = off_ < n_ ? std::memchr(p_ + off_, 0, n_ - off_) : nullptr;
        if (!z) fail(what, "unterminated string");
        const size_t len = static_cast<const uint8_t*>(z) - (p_ + off_);
        std::string s(reinterpret_cast<const char*>(p_ + off_), len);
        off_ += len + 1;
        return s;
    }
    void vec3(float* v, const char* what) {
        for (int k = 0; k < 3; k++) v[k] = f32(what);
    }
    // stream >> Transform: the matrix, then its inverse (stream/stream.h:180-184); a singular
    // matrix keeps the identity inverse of a fresh Transform.
    Transform transform(const char* what) {
        Transform t;
        for (int k = 0; k < 16; k++) t.mat.m[k] = f32(what);
        inverse(t.mat, t.inv);
        return t;
    }
    void skip(uint64_t bytes, const char* what) {
        if (bytes > n_ - off_) fail(what, "truncated");
        off_ += bytes;
    }
    [[noreturn]] void fail(const char* what, const char* why) const {
        char buf[256];
        std::snprintf(buf, sizeof(buf), "scene file: %s at byte %zu (%s)", why, off_, what);
        throw FormatError{buf};
    }

private:
    void take(void* dst, size_t bytes, const char* what) {
        if (bytes > n_ - off_) fail(what, "truncated");
        std::memcpy(dst, p_ + off_, bytes);
        off_ += bytes;
    }
    const uint8_t* p_;
    size_t n_;
    size_t off_ = 0;
};

// ---- material section -------------------------------------------------------------------------

struct Param {
    std::string name;
    int32_t channels = 0;
    float v[3] = {0, 0, 0};
};
struct Node {
    std::string name, type;
    std::vector<Param> params;
};
struct Conn {
    std::string src, src_prop, dst, dst_prop;
};
struct Graph {
    std::vector<Node> nodes;
    std::vector<Conn> conns;
};

// Shader units with their default inputs, then connections (Material::Serialize's
// parse_shader_type, material/material.cpp:160-210; the same layout in a ShaderGroupTemplate,
// material/matmanager.cpp:143-198). Inputs carry 1 float, 3 floats or (4) a global's name.
void read_graph(Reader& r, Graph& g) {
    const uint32_t nodes = r.u32("shader unit count");
    for (uint32_t i = 0; i < nodes; i++) {
        Node n;
        n.name = r.str("shader unit name");
        n.type = r.str("shader unit type");
        const uint32_t np = r.u32("shader parameter count");
        for (uint32_t j = 0; j < np; j++) {
            Param p;
            p.name = r.str("shader parameter name");
            p.channels = r.i32("shader parameter channels");
            if (p.channels == 1) {
                p.v[0] = r.f32("shader parameter value");
            } else if (p.channels == 3) {
                r.vec3(p.v, "shader parameter value");
            } else if (p.channels == 4) {
                (void)r.str("shader parameter global");
            }
            n.params.push_back(std::move(p));
        }
        g.nodes.push_back(std::move(n));
    }
    const uint32_t nc = r.u32("shader connection count");
    for (uint32_t i = 0; i < nc; i++) {
        Conn c;
        c.src = r.str("connection source");
        c.src_prop = r.str("connection source property");
        c.dst = r.str("connection target");
        c.dst_prop = r.str("connection target property");
        g.conns.push_back(std::move(c));
    }
}

struct FileMaterial {
    std::string name;
    bool surface_valid = false, volume_valid = false, transparent = false, sss = false;
    Graph surface;
};

const Param* find_param(const Node& n, const char* name) {
    for (const Param& p : n.params)
        if (p.name == name) return &p;
    return nullptr;
}

sortgpu_material default_white() {
    sortgpu_material m{};
    m.type = SORTGPU_MAT_LAMBERT;
    m.base[0] = m.base[1] = m.base[2] = 1.0f;
    return m;
}

enum class MapResult { Mapped, Approx, Unmapped };

// The GPU material for a material's surface graph: the single BXDF node that feeds the output
// node, read from its constant inputs. Diffuse (Lambert) -> Lambert (closures.cpp lambert);
// MicrofacetReflection with GGX -> GGX with conductor Fresnel eta = InteriorIOR, k =
// AbsorptionCoefficient; Glass -> rough dielectric (dielectric.cpp:32-36: IOR 1.5 inside, 1
// outside); Disney BRDF -> the Disney subset; Transparent -> the Transparent BXDF with its
// attenuation (bsdf/transparent.cpp; the material's transparent node, material.cpp:150, makes its
// shadow rays take the transparent-shadow walk). A material without a valid surface shader renders
// as white Lambert in SORT (material.cpp:230-235), so does anything else here (flagged).
// `sources`: shader unit type -> TSL source (the distribution of a microfacet node and the thin
// flag of a Disney node exist only there; one source per type, as SORT compiles it).
MapResult map_material(const FileMaterial& fm, const std::vector<std::pair<std::string, std::string>>& sources,
                       sortgpu_material& out) {
    out = default_white();
    if (!fm.surface_valid) return MapResult::Mapped;
    const Node* outn = nullptr;
    for (const Node& n : fm.surface.nodes)
        if (n.type == "SORTNodeOutput") outn = &n;
    if (!outn) return MapResult::Unmapped;
    // the exporter names the output node "ShaderOutput_" in the unit list and appends the
    // material name in connections (exporter.py:776-779)
    const Node* bxdf = nullptr;
    for (const Conn& c : fm.surface.conns)
        if (c.dst == outn->name || c.dst == outn->name + fm.name) {
            for (const Node& n : fm.surface.nodes)
                if (n.name == c.src) bxdf = &n;
        }
    if (!bxdf) return MapResult::Unmapped;
    // a transparent node anywhere but as the output's own BXDF (a mix, an add) is not carried
    bool approx = (fm.transparent && bxdf->type != "SORTNode_Material_Transparent") || fm.sss;
    for (const Conn& c : fm.surface.conns)
        if (c.dst == bxdf->name && c.dst_prop != "Normal") approx = true;  // a node drives an input
    auto f1 = [&](const char* name, float dflt) {
        const Param* p = find_param(*bxdf, name);
        return (p && p->channels == 1) ? p->v[0] : dflt;
    };
    auto f3 = [&](const char* name, float* dst, float dflt) {
        const Param* p = find_param(*bxdf, name);
        for (int k = 0; k < 3; k++) dst[k] = (p && p->channels == 3) ? p->v[k] : dflt;
    };
    const std::string& t = bxdf->type;
    const std::string* src = nullptr;
    for (const auto& kv : sources)
        if (kv.first == t) src = &kv.second;
    if (t == "SORTNode_Material_DiffuseLambert") {
        out.type = SORTGPU_MAT_LAMBERT;
        f3("Diffuse", out.base, 1.0f);
    } else if (t.rfind("SORTNode_Material_MicrofacetReflection", 0) == 0) {
        out.type = SORTGPU_MAT_GGX;
        f3("BaseColor", out.base, 1.0f);
        f3("InteriorIOR", out.eta, 0.37f);
        f3("AbsorptionCoefficient", out.k, 2.82f);
        out.roughness = f1("RoughnessU", 0.0f);
        if (f1("RoughnessV", out.roughness) != out.roughness) approx = true;
        if (!src || src->find("microfacet_reflection_ggx") == std::string::npos) approx = true;  // Blinn, Beckmann
    } else if (t == "SORTNode_Material_Glass") {
        out.type = SORTGPU_MAT_DIELECTRIC;
        f3("Reflectance", out.base, 1.0f);
        f3("Transmittance", out.trans, 1.0f);
        out.roughness = f1("RoughnessU", 0.0f);
        if (f1("RoughnessV", out.roughness) != out.roughness) approx = true;
        out.ior = 1.5f;
    } else if (t.rfind("SORTNode_Material_DisneyBRDF", 0) == 0) {
        out.type = SORTGPU_MAT_DISNEY;
        f3("BaseColor", out.base, 1.0f);
        out.metallic = f1("Metallic", 1.0f);
        out.specular = f1("Specular", 1.0f);
        out.specular_tint = f1("SpecularTint", 0.0f);
        out.roughness = f1("Roughness", 0.2f);
        out.sheen = f1("Sheen", 1.0f);
        out.sheen_tint = f1("SheenTint", 0.0f);
        out.clearcoat = f1("Clearcoat", 0.0f);
        out.clearcoat_gloss = f1("ClearcoatGlossiness", 1.0f);
        float sd[3];
        f3("ScatterDistance", sd, 0.0f);
        if (f1("Anisotropic", 0.0f) != 0.0f || f1("SpecularTransmittance", 0.0f) != 0.0f ||
            f1("DiffuseTransmittance", 0.0f) != 0.0f || sd[0] > 0.0f || sd[1] > 0.0f || sd[2] > 0.0f)
            approx = true;
        if (!src || src->find(", true ,") != std::string::npos) approx = true;  // thin surface
    } else if (t == "SORTNode_Material_Transparent") {
        out.type = SORTGPU_MAT_TRANSPARENT;
        f3("Attenuation", out.base, 1.0f);
    } else {
        return MapResult::Unmapped;
    }
    return approx ? MapResult::Approx : MapResult::Mapped;
}

}  // namespace

struct sortgpu_scene_file {
    sortgpu_scene_file_info info{};
    std::vector<float> pos, normals, uv, tangents, quad_radiance;
    std::vector<uint32_t> idx, tri_material;
    std::vector<sortgpu_quad> quads;
    std::vector<sortgpu_material> materials;
    std::vector<std::string> names;
};

namespace {

void flag(sortgpu_scene_file_info& info, int bit, uint32_t n = 1) {
    info.skipped |= 1u << bit;
    info.counts[bit] += n;
}

// ImageEvaluation::loadConfig (image_evaluation.cpp:354-378) and the integrators' Serialize.
void read_config(Reader& r, sortgpu_scene_file_info& info) {
    if (r.u32("configuration version") != 0u) r.fail("configuration version", "unsupported version");
    (void)r.str("resource path");
    info.threads = r.u32("thread count");
    info.spp = r.u32("samples per pixel");
    info.image_width = r.u32("image width");
    info.image_height = r.u32("image height");
    info.clamping = r.f32("clamping");
    info.integrator = r.u32("integrator");
    // An unknown integrator reads nothing more (MakeUniqueInstance gives null).
    const uint32_t it = info.integrator;
    const bool known = it == sid("PathTracing") || it == sid("AmbientOcclusion") || it == sid("BidirPathTracing") ||
                       it == sid("LightTracing") || it == sid("InstantRadiosity") || it == sid("DirectLight") ||
                       it == sid("WhittedRT");
    if (!known) return;
    info.max_recursive_depth = r.i32("max recursive depth");
    if (it == sid("PathTracing")) {
        info.max_bssrdf_bounces = r.i32("max bssrdf bounces");
    } else if (it == sid("AmbientOcclusion")) {
        (void)r.f32("ao max distance");
    } else if (it == sid("BidirPathTracing")) {
        (void)r.b8("bdpt mis");
    } else if (it == sid("InstantRadiosity")) {
        (void)r.i32("light path sets");
        (void)r.i32("light paths");
        (void)r.f32("min distance");
    }
}

// MatManager::ParseMatFile (matmanager.cpp:50-290).
void read_materials(Reader& r, sortgpu_scene_file& f, std::vector<FileMaterial>& mats) {
    const uint32_t nres = r.u32("resource count");
    for (uint32_t i = 0; i < nres; i++) {
        (void)r.str("resource file");
        (void)r.u32("resource type");
    }
    std::vector<std::pair<std::string, std::string>> sources;
    while (true) {
        const uint32_t tag = r.u32("material record");
        if (tag == sid("End of Material")) break;
        if (tag == sid("ShaderUnitTemplate")) {
            std::string type = r.str("shader unit type");
            std::string source = r.str("shader source");
            sources.emplace_back(std::move(type), std::move(source));
            const uint32_t nb = r.u32("shader resource count");
            for (uint32_t i = 0; i < nb; i++) {
                (void)r.str("resource handle");
                (void)r.str("resource name");
            }
        } else if (tag == sid("ShaderGroupTemplate")) {
            (void)r.str("shader group type");
            Graph g;
            read_graph(r, g);
            (void)r.str("group root shader");
            const uint32_t nout = r.u32("exposed output count");
            for (uint32_t i = 0; i < nout; i++) (void)r.str("exposed output");
            const std::string input = r.str("group input shader");
            if (!input.empty()) {
                const uint32_t nin = r.u32("exposed input count");
                for (uint32_t i = 0; i < nin; i++) (void)r.str("exposed input");
            }
        } else if (tag == sid("Material")) {
            FileMaterial m;
            m.name = r.str("material name");
            if (r.u32("surface shader tag") == sid("Surface Shader")) {
                read_graph(r, m.surface);
                m.surface_valid = true;
            }
            if (r.u32("volume shader tag") == sid("Volume Shader")) {
                Graph v;
                read_graph(r, v);
                m.volume_valid = true;
            }
            m.transparent = r.b8("has transparent node");
            m.sss = r.b8("has sss node");
            (void)r.f32("volume step");
            (void)r.u32("volume step count");
            mats.push_back(std::move(m));
        } else {
            r.fail("material record", "unknown record");
        }
    }
    for (const FileMaterial& m : mats) {
        sortgpu_material gm;
        const MapResult res = map_material(m, sources, gm);
        if (res == MapResult::Unmapped) flag(f.info, SORTGPU_FILE_UNMAPPED_MATS);
        if (res == MapResult::Approx) flag(f.info, SORTGPU_FILE_APPROX_MATS);
        if (m.volume_valid) flag(f.info, SORTGPU_FILE_VOLUMES);
        f.materials.push_back(gm);
        f.names.push_back(m.name);
    }
    f.materials.push_back(default_white());  // MatManager::GetDefaultMat (matmanager.h:54-57)
    f.names.push_back("default");
}

// Vector::Normalize (math/vector3.h:104-109): divided by its length unless that is zero.
void normalize_in_place(float* v) {
    const float len = length3(v[0], v[1], v[2]);
    if (len != 0.0f) {
        v[0] /= len;
        v[1] /= len;
        v[2] /= len;
    }
}

// Mesh::GenUV (core/mesh.cpp:60-74) for a mesh stored without texture coordinates: the direction
// from the mean vertex, as (sphericalTheta / pi, sphericalPhi / 2pi) (bxdf_utils.h:80-87, the
// constants of math/utils.h:27-33).
void gen_uv(sortgpu_scene_file& f, uint32_t base, uint32_t nv) {
    if (nv == 0) return;
    const float* p = f.pos.data() + 3 * (size_t)base;
    float c[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t i = 0; i < nv; i++)
        for (int k = 0; k < 3; k++) c[k] = c[k] + p[3 * (size_t)i + k];
    for (int k = 0; k < 3; k++) c[k] /= (float)nv;
    float* uv = f.uv.data() + 2 * (size_t)base;
    for (uint32_t i = 0; i < nv; i++) {
        float d[3] = {p[3 * (size_t)i] - c[0], p[3 * (size_t)i + 1] - c[1], p[3 * (size_t)i + 2] - c[2]};
        normalize_in_place(d);
        const float theta = acosf(std::min(std::max(d[1], -1.0f), 1.0f));
        float phi = atan2f(d[2], d[0]);
        if (phi < 0.0f) phi = phi + 2.0f * 3.1415926f;
        uv[2 * (size_t)i] = theta * 0.3183099f;
        uv[2 * (size_t)i + 1] = phi * 0.15915494f;
    }
}

// Mesh::genTagentForTri (core/mesh.cpp:76-108): dp/du of the triangle from its texture
// coordinates, or, when they are degenerate, coordinateSystem's first axis (math/vector3.h:171-180)
// about the face normal.
void tri_tangent(const float* p0, const float* p1, const float* p2, const float* t0, const float* t1,
                 const float* t2, float* out) {
    const float du1 = t0[0] - t2[0], du2 = t1[0] - t2[0];
    const float dv1 = t0[1] - t2[1], dv2 = t1[1] - t2[1];
    const float det = du1 * dv2 - dv1 * du2;
    if (det == 0.0f) {
        const float a[3] = {p0[0] - p1[0], p0[1] - p1[1], p0[2] - p1[2]};
        const float b[3] = {p0[0] - p2[0], p0[1] - p2[1], p0[2] - p2[2]};
        float n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        normalize_in_place(n);  // normalize() (vector3.h:163-168) gives the same floats
        if (std::fabs(n[0]) > std::fabs(n[1])) {
            const float inv = 1.0f / std::sqrt(n[0] * n[0] + n[2] * n[2]);
            out[0] = -n[2] * inv;
            out[1] = 0.0f;
            out[2] = n[0] * inv;
        } else {
            const float inv = 1.0f / std::sqrt(n[1] * n[1] + n[2] * n[2]);
            out[0] = 0.0f;
            out[1] = n[2] * inv;
            out[2] = -n[1] * inv;
        }
        return;
    }
    for (int k = 0; k < 3; k++) {
        const float dp1 = p0[k] - p2[k], dp2 = p1[k] - p2[k];
        out[k] = (dv2 * dp1 - dv1 * dp2) / det;
    }
}

// Mesh::GenSmoothTagent (core/mesh.cpp:42-58): each vertex's tangent is the normalised sum, in
// triangle order, of the tangents of the triangles that use it (world-space positions; the
// transform is not applied to tangents).
void gen_smooth_tangents(sortgpu_scene_file& f, uint32_t base, uint32_t nv, uint32_t first_tri) {
    const size_t t0 = 3 * (size_t)base;
    f.tangents.resize(t0 + 3 * (size_t)nv, 0.0f);
    float* tan = f.tangents.data();
    const float* pos = f.pos.data();
    const float* uv = f.uv.data();
    const size_t ntri = f.idx.size() / 3;
    for (size_t t = first_tri; t < ntri; t++) {
        const uint32_t* v = f.idx.data() + 3 * t;
        float tt[3];
        tri_tangent(pos + 3 * (size_t)v[0], pos + 3 * (size_t)v[1], pos + 3 * (size_t)v[2], uv + 2 * (size_t)v[0],
                    uv + 2 * (size_t)v[1], uv + 2 * (size_t)v[2], tt);
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 3; k++) tan[3 * (size_t)v[c] + k] += tt[k];
    }
    for (uint32_t i = 0; i < nv; i++) normalize_in_place(tan + t0 + 3 * (size_t)i);
}

// Mesh::Serialize (core/mesh.cpp:111-180), then MeshVisual::ApplyTransform (entity/visual.cpp:35-39):
// Mesh::ApplyTransform (mesh.cpp:26-40), Mesh::GenUV, Mesh::GenSmoothTagent.
void read_mesh(Reader& r, sortgpu_scene_file& f, const Transform& tr, uint32_t nmats) {
    const bool has_uv = r.b8("mesh has uv");
    const uint32_t nv = r.u32("vertex count");
    if ((uint64_t)nv * 32u > r.remaining()) r.fail("vertex count", "truncated");
    f.pos.reserve(f.pos.size() + 3 * (size_t)nv);
    f.normals.reserve(f.normals.size() + 3 * (size_t)nv);
    f.uv.reserve(f.uv.size() + 2 * (size_t)nv);
    const uint32_t base = (uint32_t)(f.pos.size() / 3);
    const float* m = tr.mat.m;
    const float* iv = tr.inv.m;
    for (uint32_t i = 0; i < nv; i++) {
        float p[3], n[3], t[2];
        r.vec3(p, "vertex position");
        r.vec3(n, "vertex normal");
        t[0] = r.f32("vertex uv");
        t[1] = r.f32("vertex uv");
        // Matrix::TransformPoint (math/matrix.cpp:70-82)
        float x = p[0] * m[0] + p[1] * m[1] + p[2] * m[2] + m[3];
        float y = p[0] * m[4] + p[1] * m[5] + p[2] * m[6] + m[7];
        float z = p[0] * m[8] + p[1] * m[9] + p[2] * m[10] + m[11];
        const float w = p[0] * m[12] + p[1] * m[13] + p[2] * m[14] + m[15];
        if (w != 1.0f) {
            x = x / w;
            y = y / w;
            z = z / w;
        }
        f.pos.insert(f.pos.end(), {x, y, z});
        // Transform::TransformNormal of Vector::Normalize (transform.h:54-56, vector3.h:104-109):
        // the transposed inverse applied as a vector, no renormalisation afterwards.
        const float len = length3(n[0], n[1], n[2]);
        if (len != 0.0f) {
            n[0] /= len;
            n[1] /= len;
            n[2] /= len;
        }
        const float nx = n[0] * iv[0] + n[1] * iv[4] + n[2] * iv[8];
        const float ny = n[0] * iv[1] + n[1] * iv[5] + n[2] * iv[9];
        const float nz = n[0] * iv[2] + n[1] * iv[6] + n[2] * iv[10];
        f.normals.insert(f.normals.end(), {nx, ny, nz});
        f.uv.insert(f.uv.end(), {t[0], t[1]});
    }
    const uint32_t nt = r.u32("triangle count");
    if ((uint64_t)nt * 16u > r.remaining()) r.fail("triangle count", "truncated");
    f.idx.reserve(f.idx.size() + 3 * (size_t)nt);
    f.tri_material.reserve(f.tri_material.size() + nt);
    for (uint32_t i = 0; i < nt; i++) {
        uint32_t v[3];
        for (int k = 0; k < 3; k++) {
            v[k] = r.u32("triangle index");
            if (v[k] >= nv) r.fail("triangle index", "vertex index out of range");
        }
        const int32_t mat = r.i32("triangle material");
        f.idx.insert(f.idx.end(), {base + v[0], base + v[1], base + v[2]});
        // MatManager::GetMaterial: an id outside the material list is the default material
        f.tri_material.push_back((mat >= 0 && (uint32_t)mat < nmats) ? (uint32_t)mat : nmats);
    }
    const uint32_t vol = r.u32("volume tag");
    if (vol == sid("has_volume")) {
        // MediumDensity::Serialize (medium/mediumdata.cpp:26-38); MediumColor reads nothing
        const uint32_t w = r.u32("volume width"), h = r.u32("volume height"), d = r.u32("volume depth");
        if (w && h && d) r.skip((uint64_t)w * h * d * 4u, "volume density");
        flag(f.info, SORTGPU_FILE_VOLUMES);
    } else if (vol != sid("no_volume")) {
        r.fail("volume tag", "unknown volume tag");
    }
    if (r.u32("end of mesh") != sid("end of mesh")) r.fail("end of mesh", "missing end-of-mesh tag");
    if (!has_uv) gen_uv(f, base, nv);
    gen_smooth_tangents(f, base, nv, (uint32_t)(f.idx.size() / 3) - nt);
}

// HairVisual::Serialize (entity/visual.cpp:44-100): read and counted, not carried.
void read_hair(Reader& r, sortgpu_scene_file& f) {
    const uint32_t n = r.u32("hair count");
    (void)r.f32("hair tip width");
    (void)r.f32("hair bottom width");
    (void)r.i32("hair material");
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t steps = r.u32("hair steps");
        if (steps == 0) continue;
        r.skip((uint64_t)(steps + 1u) * 12u, "hair points");
    }
    flag(f.info, SORTGPU_FILE_HAIR, n);
}

// AreaLightEntity::Serialize (light_entity.cpp:99-155): the transform loses its scale (the
// lengths of its first three columns) into the shape, the radiance takes Blender's power factor.
void read_area_light(Reader& r, sortgpu_scene_file& f) {
    const Transform t = r.transform("area light transform");
    const float* m = t.mat.m;
    const float sx = length3(m[0], m[4], m[8]);
    const float sy = length3(m[1], m[5], m[9]);
    const float sz = length3(m[2], m[6], m[10]);
    Mat4 down = identity(), up = identity();  // Scale(1/sx, 1/sy, 1/sz).matrix, Scale(sx, sy, sz).matrix
    down.m[0] = 1 / sx;
    down.m[5] = 1 / sy;
    down.m[10] = 1 / sz;
    up.m[0] = sx;
    up.m[5] = sy;
    up.m[10] = sz;
    const Mat4 mat = mul(t.mat, down);
    const Mat4 inv = mul(up, t.inv);
    const float energy = r.f32("light energy");
    float rgb[3];
    r.vec3(rgb, "light color");
    const uint32_t shape = r.u32("area light shape");
    float scale;
    sortgpu_quad q{};
    if (shape == sid("SQUARE")) {
        const float size = r.f32("area light size");
        q.size_x = q.size_y = size;
        scale = 0.8f * energy / (size * size * sx * sy * kPi);
    } else if (shape == sid("RECTANGLE")) {
        const float a = r.f32("area light size x");
        const float b = r.f32("area light size y");
        q.size_x = a;
        q.size_y = b;
        scale = 0.8f * energy / (a * sx * b * sy * kPi);
    } else if (shape == sid("DISK")) {
        (void)r.f32("area light radius");
        flag(f.info, SORTGPU_FILE_DISK_LIGHTS);
        return;
    } else {
        return;  // an unrecognised shape is only warned about by SORT
    }
    std::memcpy(q.mat, mat.m, sizeof(q.mat));
    std::memcpy(q.inv, inv.m, sizeof(q.inv));
    f.quads.push_back(q);
    f.quad_radiance.insert(f.quad_radiance.end(), {rgb[0] * scale, rgb[1] * scale, rgb[2] * scale});
}

// PerspectiveCameraEntity::Serialize (camera_entity.cpp:21-32) and the field of view that
// PerspectiveCamera::PreProcess (camera/perspective.cpp:26-63) gives the vertical axis.
void read_camera(Reader& r, sortgpu_scene_file_info& info) {
    r.vec3(info.eye, "camera eye");
    r.vec3(info.up, "camera up");
    r.vec3(info.target, "camera target");
    info.camera_width = r.u32("camera width");
    info.camera_height = r.u32("camera height");
    info.lens_radius = r.f32("lens radius");
    info.sensor_w = r.f32("sensor width");
    info.sensor_h = r.f32("sensor height");
    info.aspect_fit = r.i32("aspect fit");
    info.pixel_aspect_w = r.f32("pixel aspect x");
    info.pixel_aspect_h = r.f32("pixel aspect y");
    info.fov = r.f32("field of view");
    info.has_camera = 1;
    const float w = (float)info.camera_width, h = (float)info.camera_height;
    const float aspect = w / h * info.pixel_aspect_w / info.pixel_aspect_h;
    float y_scale = 1.0f / std::tan(info.fov * 0.5f);
    if (info.sensor_w != 0.0f && info.sensor_h != 0.0f) {
        if (info.aspect_fit == 1 || (info.aspect_fit != 2 && aspect > 1)) y_scale = y_scale * aspect;
    }
    info.camera_fov_y = 2.0f * std::atan(1.0f / y_scale);
    if (info.lens_radius != 0.0f || info.pixel_aspect_w != info.pixel_aspect_h) flag(info, SORTGPU_FILE_CAMERA);
}

void read_scene(Reader& r, sortgpu_scene_file& f, uint32_t nmats) {
    sortgpu_scene_file_info& info = f.info;
    if (r.u32("verification bits") != sid("verification bits")) r.fail("verification bits", "not a scene stream");
    bool any_light = false;
    bool have_sky = false;
    while (true) {
        const uint32_t cls = r.u32("entity class");
        if (cls == sid("End of Entities")) break;
        if (cls == sid("PerspectiveCameraEntity")) {
            read_camera(r, info);
        } else if (cls == sid("VisualEntity")) {
            const Transform t = r.transform("visual transform");
            const uint32_t nvis = r.u32("visual count");
            for (uint32_t i = 0; i < nvis; i++) {
                const uint32_t vc = r.u32("visual class");
                if (vc == sid("MeshVisual")) read_mesh(r, f, t, nmats);
                else if (vc == sid("HairVisual")) read_hair(r, f);
                else r.fail("visual class", "unknown visual");
            }
        } else if (cls == sid("PointLightEntity") || cls == sid("DirLightEntity") || cls == sid("SpotLightEntity")) {
            (void)r.transform("light transform");
            (void)r.f32("light energy");
            float rgb[3];
            r.vec3(rgb, "light color");
            if (cls == sid("SpotLightEntity")) {
                (void)r.f32("spot falloff start");
                (void)r.f32("spot falloff range");
            }
            any_light = true;
            flag(info, SORTGPU_FILE_PUNCTUAL_LIGHTS);
        } else if (cls == sid("SkyLightEntity")) {
            // SkyLightEntity::Serialize (light_entity.cpp:74-92)
            (void)r.transform("sky transform");
            const float intensity = r.f32("sky intensity");
            float rgb[3];
            r.vec3(rgb, "sky color");
            const std::string image = r.str("sky image");
            for (int k = 0; k < 3; k++) info.env_radiance[k] = rgb[k] * intensity;
            if (!image.empty()) flag(info, SORTGPU_FILE_SKY_IMAGE);
            any_light = have_sky = true;
        } else if (cls == sid("AreaLightEntity")) {
            read_area_light(r, f);
            any_light = true;
        } else {
            r.fail("entity class", "unknown entity");
        }
    }
    // Scene::LoadScene adds a default sky of radiance 0.5 when no light was read (scene.cpp:135-141)
    if (!any_light && !have_sky) info.env_radiance[0] = info.env_radiance[1] = info.env_radiance[2] = 0.5f;
    // the accelerator record (scene.cpp:160-177; Serialize of each accelerator class)
    info.accelerator = r.u32("accelerator");
    const uint32_t a = info.accelerator;
    if (a == sid("Bvh") || a == sid("KDTree") || a == sid("OcTree") || a == sid("Qbvh") || a == sid("Obvh")) {
        info.accel_max_depth = r.i32("accelerator max depth");
        info.accel_max_leaf = r.i32("accelerator max leaf");
    }
}

void finish(sortgpu_scene_file& f) {
    sortgpu_scene_file_info& i = f.info;
    i.nverts = (uint32_t)(f.pos.size() / 3);
    i.ntris = (uint32_t)(f.idx.size() / 3);
    i.nquads = (uint32_t)f.quads.size();
    i.nmaterials = (uint32_t)f.materials.size();
    i.pos = f.pos.data();
    i.normals = f.normals.data();
    i.uv = f.uv.data();
    i.tangents = f.tangents.data();
    i.idx = f.idx.data();
    i.tri_material = f.tri_material.data();
    i.quads = f.quads.data();
    i.quad_radiance = f.quad_radiance.data();
    i.materials = f.materials.data();
}

int parse(const uint8_t* data, size_t size, sortgpu_scene_file** out) {
    std::unique_ptr<sortgpu_scene_file> f(new (std::nothrow) sortgpu_scene_file());
    if (!f) {
        sortgpu::set_error(nullptr, "scene file: host allocation failed");
        return SORTGPU_ERR_OOM;
    }
    try {
        Reader r(data, size);
        read_config(r, f->info);
        std::vector<FileMaterial> mats;
        read_materials(r, *f, mats);
        read_scene(r, *f, (uint32_t)mats.size());
        if (f->idx.size() / 3 > 0xFFFFFFFFull / 3) r.fail("triangles", "too many triangles");
    } catch (const FormatError& e) {
        sortgpu::set_error(nullptr, "%s", e.msg.c_str());
        return SORTGPU_ERR_FORMAT;
    } catch (const std::bad_alloc&) {
        sortgpu::set_error(nullptr, "scene file: host allocation failed");
        return SORTGPU_ERR_OOM;
    }
    finish(*f);
    *out = f.release();
    return SORTGPU_OK;
}

}  // namespace

extern "C" {

int sortgpu_scene_file_parse(const void* data, uint64_t size, sortgpu_scene_file** out) {
    if (!out || (!data && size)) {
        sortgpu::set_error(nullptr, "scene file: null argument");
        return SORTGPU_ERR_ARG;
    }
    *out = nullptr;
    return parse(static_cast<const uint8_t*>(data), (size_t)size, out);
}

int sortgpu_scene_file_load(const char* path, sortgpu_scene_file** out) {
    if (!path || !out) {
        sortgpu::set_error(nullptr, "scene file: null argument");
        return SORTGPU_ERR_ARG;
    }
    *out = nullptr;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) {
        sortgpu::set_error(nullptr, "scene file: cannot open %s", path);
        return SORTGPU_ERR_FORMAT;
    }
    std::vector<uint8_t> buf;
    int rc = SORTGPU_OK;
    try {
        uint8_t chunk[1 << 16];
        size_t got;
        while ((got = std::fread(chunk, 1, sizeof(chunk), fp)) > 0) buf.insert(buf.end(), chunk, chunk + got);
        if (std::ferror(fp)) {
            sortgpu::set_error(nullptr, "scene file: read error on %s", path);
            rc = SORTGPU_ERR_FORMAT;
        }
    } catch (const std::bad_alloc&) {
        sortgpu::set_error(nullptr, "scene file: host allocation failed");
        rc = SORTGPU_ERR_OOM;
    }
    std::fclose(fp);
    if (rc != SORTGPU_OK) return rc;
    return parse(buf.data(), buf.size(), out);
}

const sortgpu_scene_file_info* sortgpu_scene_file_get(const sortgpu_scene_file* file) {
    return file ? &file->info : nullptr;
}

const char* sortgpu_scene_file_material_name(const sortgpu_scene_file* file, uint32_t i) {
    if (!file || i >= file->names.size()) return nullptr;
    return file->names[i].c_str();
}

void sortgpu_scene_file_destroy(sortgpu_scene_file* file) { delete file; }

}  // extern "C"
