// Host-side helpers of the path tracer: device allocation lists, the path workspace, the
// request-table ring, scene images and the camera basis.
#include "pt_host.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>

namespace sortgpu {

int DeviceAllocs::alloc(sortgpu_ctx* ctx, void** p, size_t n, const char* what) {
    *p = nullptr;
    hipError_t e = hipMalloc(p, std::max<size_t>(n, 16));
    if (e != hipSuccess) {
        *p = nullptr;
        return e == hipErrorOutOfMemory ? (set_error(ctx, "%s: out of device memory", what), SORTGPU_ERR_OOM)
                                        : hip_fail(ctx, e, what);
    }
    ptrs.push_back(*p);
    bytes += n;
    return SORTGPU_OK;
}

int DeviceAllocs::upload(sortgpu_ctx* ctx, const void* src, size_t n, void** dst, const char* what) {
    *dst = nullptr;
    if (!src || !n) return SORTGPU_OK;
    int rc = alloc(ctx, dst, n, what);
    if (rc) return rc;
    hipError_t e = hipMemcpy(*dst, src, n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(ctx, e, what);
    return SORTGPU_OK;
}

void DeviceAllocs::free_all() {
    for (void* p : ptrs) hipFree(p);
    ptrs.clear();
    bytes = 0;
}

namespace {

template <class T>
int ws(DeviceAllocs& m, sortgpu_ctx* ctx, T** out, size_t count) {
    void* p = nullptr;
    int rc = m.alloc(ctx, &p, count * sizeof(T), "hipMalloc(path workspace)");
    *out = static_cast<T*>(p);
    return rc;
}

float lum(const float* c) { return 0.212671f * c[0] + 0.715160f * c[1] + 0.072169f * c[2]; }

}  // namespace

size_t PathWorkspace::bytes_for(uint32_t paths) {
    return (size_t)std::max(paths, 1u) * kPathBytes + 16 * 4 + 32 + ((size_t)kSortBlocks + 1) * kSortBuckets * 4;
}

int PathWorkspace::ensure(sortgpu_ctx* ctx, uint32_t paths) {
    if (paths <= cap) return SORTGPU_OK;
    release();
    paths = std::max(paths, 1u);
    const size_t n = paths;
    int rc;
    if ((rc = ws(mem, ctx, &ps.rec, 4 * n)) || (rc = ws(mem, ctx, &q.ray[0], 2 * n)) ||
        (rc = ws(mem, ctx, &q.ray[1], 2 * n)) || (rc = ws(mem, ctx, &q.pid[0], n)) ||
        (rc = ws(mem, ctx, &q.pid[1], n)) || (rc = ws(mem, ctx, &q.rsc[0], n)) || (rc = ws(mem, ctx, &q.rsc[1], n)) ||
        (rc = ws(mem, ctx, &q.hit, n)) || (rc = ws(mem, ctx, &q.sray, 4 * n)) || (rc = ws(mem, ctx, &q.ssid, 2 * n)) ||
        (rc = ws(mem, ctx, &q.ssc, 2 * n)) ||
        (rc = ws(mem, ctx, &q.count, 16)) || (rc = ws(mem, ctx, &q.totals, 4)) || (rc = ws(mem, ctx, &q.skey, n)) ||
        (rc = ws(mem, ctx, &q.sperm, n)) || (rc = ws(mem, ctx, &q.shist, (size_t)kSortBlocks * kSortBuckets)) ||
        (rc = ws(mem, ctx, &q.sbase, kSortBuckets))) {
        release();
        return rc;
    }
    cap = paths;
    return SORTGPU_OK;
}

void PathWorkspace::release() {
    mem.free_all();
    ps = PathState{};
    q = Queues{};
    cap = 0;
}

int ReqRing::init(sortgpu_ctx* ctx) {
    const size_t n = (size_t)kSlots * sizeof(WaveDesc);
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&host), n, hipHostMallocDefault);
    if (e != hipSuccess) {
        host = nullptr;
        return hip_fail(ctx, e, "hipHostMalloc(request ring)");
    }
    if ((e = hipMalloc(reinterpret_cast<void**>(&dev), n)) != hipSuccess) {
        dev = nullptr;
        release();
        return hip_fail(ctx, e, "hipMalloc(request ring)");
    }
    for (int i = 0; i < kSlots; i++)
        if ((e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming)) != hipSuccess) {
            ev[i] = nullptr;
            release();
            return hip_fail(ctx, e, "hipEventCreate(request ring)");
        }
    return SORTGPU_OK;
}

void ReqRing::release() {
    for (int i = 0; i < kSlots; i++)
        if (ev[i]) {
            hipEventSynchronize(ev[i]);
            hipEventDestroy(ev[i]);
            ev[i] = nullptr;
        }
    if (host) hipHostFree(host);
    if (dev) hipFree(dev);
    host = nullptr;
    dev = nullptr;
}

int ReqRing::push(sortgpu_ctx* ctx, const WaveDesc& w, hipStream_t s, WaveDesc* dst, const WaveDesc** d_desc) {
    const uint32_t n = w.ct.n;
    if (n == 0 || n > (uint32_t)kMaxChunks) {
        set_error(ctx, "wave descriptor: %u requests (1..%d)", n, kMaxChunks);
        return SORTGPU_ERR_ARG;
    }
    const uint32_t slot = next;
    next = (next + 1) % kSlots;
    hipError_t e = hipEventSynchronize(ev[slot]);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventSynchronize(wave ring)");
    WaveDesc* h = host + slot;
    WaveDesc* d = dst ? dst : dev + slot;
    const size_t bytes = offsetof(WaveDesc, reqs) + n * sizeof(ReqDev);
    std::memcpy(h, &w, bytes);
    if ((e = hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail(ctx, e, "hipMemcpyAsync(wave descriptor)");
    last = slot;
    *d_desc = d;
    return SORTGPU_OK;
}

int ReqRing::done(sortgpu_ctx* ctx, hipStream_t s) {
    hipError_t e = hipEventRecord(ev[last], s);
    return e == hipSuccess ? SORTGPU_OK : hip_fail(ctx, e, "hipEventRecord(request ring)");
}

int prepare_scene(sortgpu_ctx* ctx, const sortgpu_accel* acc, const sortgpu_scene_desc* desc, bool with_tree,
                  SceneImage& im) {
    if (!acc || !desc) {
        set_error(ctx, "scene: null argument");
        return SORTGPU_ERR_ARG;
    }
    const uint32_t ntris = (uint32_t)(acc->idx.size() / 3), nverts = (uint32_t)(acc->pos.size() / 3);
    const uint32_t nquads = (uint32_t)acc->quads.size();
    if (desc->nmaterials && !desc->materials) {
        set_error(ctx, "scene: materials is NULL");
        return SORTGPU_ERR_ARG;
    }
    if (!(desc->clamp >= 0.0f)) {
        set_error(ctx, "scene: clamp must be >= 0 (0 = none)");
        return SORTGPU_ERR_ARG;
    }
    if (nquads && !desc->quad_radiance) {
        set_error(ctx, "scene: quad_radiance is NULL but the accelerator has quads");
        return SORTGPU_ERR_ARG;
    }
    if (desc->tri_material) {
        for (uint32_t i = 0; i < ntris; i++)
            if (desc->tri_material[i] >= std::max(desc->nmaterials, 1u)) {
                set_error(ctx, "scene: triangle %u has material %u (of %u)", i, desc->tri_material[i],
                          desc->nmaterials);
                return SORTGPU_ERR_ARG;
            }
    } else if (ntris && desc->nmaterials == 0) {
        set_error(ctx, "scene: triangles need at least one material");
        return SORTGPU_ERR_ARG;
    }
    for (uint32_t m = 0; m < desc->nmaterials; m++)
        if (desc->materials[m].type > MAT_TRANSPARENT) {
            set_error(ctx, "scene: material %u has unknown type %u", m, desc->materials[m].type);
            return SORTGPU_ERR_ARG;
        }
    try {
        im = SceneImage{};
        im.with_tree = with_tree;
        im.tree = acc->view;
        im.tree.nodes = nullptr;
        im.tree.prims = nullptr;
        im.tree.quads = nullptr;
        if (with_tree) {
            im.nodes = acc->host.node_bytes;
            im.prims = acc->host.prims;
            im.quads = acc->host.quads;
        }
        im.idx = acc->idx;
        im.pos = acc->pos;
        if (desc->normals) im.nrm.assign(desc->normals, desc->normals + 3 * (size_t)nverts);
        if (desc->tangents) im.tan.assign(desc->tangents, desc->tangents + 3 * (size_t)nverts);
        if (desc->tri_material) im.tri_mat.assign(desc->tri_material, desc->tri_material + ntris);

        // materials + the default white Lambert
        im.mats.resize(desc->nmaterials + 1);
        std::memset(im.mats.data(), 0, im.mats.size() * sizeof(MatRec));
        for (uint32_t m = 0; m < desc->nmaterials; m++) {
            const sortgpu_material& s = desc->materials[m];
            MatRec& d = im.mats[m];
            d.type = s.type;
            d.roughness = s.roughness;
            d.ior = s.ior > 0.0f ? s.ior : 1.5f;
            d.base = v3(s.base[0], s.base[1], s.base[2]);
            d.trans = v3(s.trans[0], s.trans[1], s.trans[2]);
            d.eta = v3(s.eta[0], s.eta[1], s.eta[2]);
            d.k = v3(s.k[0], s.k[1], s.k[2]);
            d.metallic = s.metallic;
            d.specular = s.specular;
            d.specular_tint = s.specular_tint;
            d.sheen = s.sheen;
            d.sheen_tint = s.sheen_tint;
            d.clearcoat = s.clearcoat;
            d.clearcoat_gloss = s.clearcoat_gloss;
        }
        im.mats.back().type = MAT_LAMBERT;
        im.mats.back().base = v3(1.f, 1.f, 1.f);

        // transparency per primitive id (Material::EvaluateTransparency, material/material.h:159-165):
        // a Transparent material's attenuation; other materials, and the quads' default white
        // Lambert, have no transparent node
        bool any_transparent = false;
        for (uint32_t m = 0; m < desc->nmaterials; m++) any_transparent |= desc->materials[m].type == MAT_TRANSPARENT;
        if (any_transparent) {
            im.transp.assign(4 * ((size_t)ntris + nquads), 0.0f);
            for (uint32_t t = 0; t < ntris; t++) {
                const MatRec& mr = im.mats[desc->tri_material ? desc->tri_material[t] : 0];
                if (mr.type != MAT_TRANSPARENT) continue;
                im.transp[4 * (size_t)t] = mr.base.x;
                im.transp[4 * (size_t)t + 1] = mr.base.y;
                im.transp[4 * (size_t)t + 2] = mr.base.z;
                im.transp[4 * (size_t)t + 3] = 1.0f;
            }
        }

        // lights: area quads in quad order, then the environment (Scene::genLightDistribution)
        std::vector<float> power;
        im.quad_light.resize(nquads);
        for (uint32_t qi = 0; qi < nquads; qi++) {
            const sortgpu_quad& q = acc->quads[qi];
            LightRec l;
            std::memset(&l, 0, sizeof(l));
            std::memcpy(l.mat, q.mat, 12 * sizeof(float));
            std::memcpy(l.inv, q.inv, 16 * sizeof(float));
            const float* Lq = desc->quad_radiance + 3 * qi;
            l.L = v3(Lq[0], Lq[1], Lq[2]);
            l.area = q.size_x * q.size_y;
            l.half_x = q.size_x * 0.5f;
            l.half_y = q.size_y * 0.5f;
            l.prim = ntris + qi;
            l.kind = 0;
            im.quad_light[qi] = (int32_t)im.lights.size();
            im.lights.push_back(l);
            power.push_back(l.area * lum(Lq) * 6.28318530717958647692f);  // AreaLight::Power
        }
        int32_t env_idx = -1;
        const float* E = desc->env_radiance;
        if (E[0] != 0.0f || E[1] != 0.0f || E[2] != 0.0f) {
            LightRec l;
            std::memset(&l, 0, sizeof(l));
            l.L = v3(E[0], E[1], E[2]);
            l.kind = 1;
            env_idx = (int32_t)im.lights.size();
            im.lights.push_back(l);
            const float* b = acc->host.bbox;
            const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
            const float r = std::sqrt(dx * dx + dy * dy + dz * dz) * 0.5f;
            power.push_back(r * r * 3.14159265358979323846f * lum(E));  // SkyLight::Power
        }
        const uint32_t nl = (uint32_t)im.lights.size();
        im.cdf.assign(nl + 1, 0.0f);  // Distribution1D (core/samplemethod.h:132-152)
        for (uint32_t i = 0; i < nl; i++) im.cdf[i + 1] = im.cdf[i] + power[i];
        const float sum = nl ? im.cdf[nl] : 0.0f;
        for (uint32_t i = 0; i <= nl && nl; i++) im.cdf[i] = sum != 0.0f ? im.cdf[i] / sum : (float)i / (float)nl;

        SceneDev& h = im.dev;
        std::memset(&h, 0, sizeof(h));
        h.ntris = ntris;
        h.nquads = nquads;
        h.nlights = nl;
        h.nmats = desc->nmaterials;
        h.env_light = env_idx;
        h.env_L = v3(E[0], E[1], E[2]);
        h.clamp = desc->clamp;
    } catch (const std::bad_alloc&) {
        set_error(ctx, "scene: host allocation failed");
        return SORTGPU_ERR_OOM;
    }
    return SORTGPU_OK;
}

namespace {
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

template <class T>
size_t vbytes(const std::vector<T>& v) {
    return v.size() * sizeof(T);
}

// Walks the arrays of an image in block order, calling f(host data, bytes, device pointer slot).
template <class F>
void for_each_array(const SceneImage& im, SceneDev& d, F f) {
    if (im.with_tree) {
        f(im.nodes.data(), vbytes(im.nodes), reinterpret_cast<const void**>(&d.tree.nodes));
        f(im.prims.data(), vbytes(im.prims), reinterpret_cast<const void**>(&d.tree.prims));
        f(im.quads.data(), vbytes(im.quads), reinterpret_cast<const void**>(&d.tree.quads));
    }
    f(im.idx.data(), vbytes(im.idx), reinterpret_cast<const void**>(&d.idx));
    f(im.pos.data(), vbytes(im.pos), reinterpret_cast<const void**>(&d.pos));
    f(im.nrm.data(), vbytes(im.nrm), reinterpret_cast<const void**>(&d.nrm));
    f(im.tan.data(), vbytes(im.tan), reinterpret_cast<const void**>(&d.tan));
    f(im.tri_mat.data(), vbytes(im.tri_mat), reinterpret_cast<const void**>(&d.tri_mat));
    f(im.mats.data(), vbytes(im.mats), reinterpret_cast<const void**>(&d.mats));
    f(im.lights.data(), vbytes(im.lights), reinterpret_cast<const void**>(&d.lights));
    f(im.cdf.data(), vbytes(im.cdf), reinterpret_cast<const void**>(&d.light_cdf));
    f(im.quad_light.data(), vbytes(im.quad_light), reinterpret_cast<const void**>(&d.quad_light));
    f(im.transp.data(), vbytes(im.transp), reinterpret_cast<const void**>(&d.transp));
}
}  // namespace

size_t SceneImage::block_bytes() const {
    size_t total = 0;
    SceneDev d{};
    for_each_array(*this, d, [&](const void*, size_t n, const void**) { total += align256(n); });
    return std::max<size_t>(total, 256);
}

int SceneImage::place(sortgpu_ctx* ctx, void* block, SceneDev& out, hipStream_t s) const {
    out = dev;
    out.tree = tree;
    size_t off = 0;
    hipError_t e = hipSuccess;
    char* base = static_cast<char*>(block);
    for_each_array(*this, out, [&](const void* src, size_t n, const void** slot) {
        *slot = nullptr;
        if (!n) return;
        *slot = base + off;
        if (e == hipSuccess) e = hipMemcpyAsync(base + off, src, n, hipMemcpyHostToDevice, s);
        off += align256(n);
    });
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync(scene)");
    return SORTGPU_OK;
}

int make_camera(sortgpu_ctx* ctx, const sortgpu_camera* cam, uint32_t width, uint32_t height, CamDev& cd) {
    const V3 eye = v3(cam->eye[0], cam->eye[1], cam->eye[2]);
    V3 f = v3(cam->target[0] - eye.x, cam->target[1] - eye.y, cam->target[2] - eye.z);
    const float fl = std::sqrt(f.x * f.x + f.y * f.y + f.z * f.z);
    if (!(fl > 0.0f) || !(cam->fov_y > 0.0f) || !(cam->fov_y < 3.14159f)) {
        set_error(ctx, "camera: degenerate (eye == target or fov outside (0, pi))");
        return SORTGPU_ERR_ARG;
    }
    f = f / fl;
    const V3 up = v3(cam->up[0], cam->up[1], cam->up[2]);
    V3 r = v3(up.y * f.z - up.z * f.y, up.z * f.x - up.x * f.z, up.x * f.y - up.y * f.x);  // cross(up, dir)
    const float rl = std::sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
    if (!(rl > 0.0f)) {
        set_error(ctx, "camera: up is parallel to the view direction");
        return SORTGPU_ERR_ARG;
    }
    r = r / rl;
    const V3 u2 = v3(f.y * r.z - f.z * r.y, f.z * r.x - f.x * r.z, f.x * r.y - f.y * r.x);  // cross(dir, right)
    cd.eye = eye;
    cd.right = r;
    cd.up = u2;
    cd.fwd = f;
    const float ys = 1.0f / std::tan(cam->fov_y * 0.5f);
    cd.y_scale = ys;
    cd.x_scale = ys / ((float)width / (float)height);
    cd.width = width;
    cd.height = height;
    return SORTGPU_OK;
}

}  // namespace sortgpu
