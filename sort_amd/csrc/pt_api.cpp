// C ABI of the path tracer: scene upload, render, batched Li, statistics.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/sortgpu.h"
#include "api_internal.h"
#include "pt_launch.h"

using namespace sortgpu;

struct sortgpu_scene {
    sortgpu_accel* accel = nullptr;
    SceneDev host{};                 // pointers are device addresses
    SceneDev* d_scene = nullptr;
    std::vector<void*> allocs;       // device allocations owned by the scene
    // workspace (grown on demand)
    uint32_t cap_paths = 0, cap_pix = 0;
    PathState ps{};
    Queues q{};
    float4* film = nullptr;
    std::vector<void*> ws;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t last_paths = 0;
};

namespace {

template <class T>
int upload(sortgpu_scene* sc, const T* src, size_t count, const T** dst, const char* what) {
    *dst = nullptr;
    if (!src || !count) return SORTGPU_OK;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) return hip_fail(sc->accel->ctx, e, what);
    sc->allocs.push_back(p);
    e = hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(sc->accel->ctx, e, what);
    *dst = static_cast<const T*>(p);
    return SORTGPU_OK;
}

void free_ws(sortgpu_scene* sc) {
    for (void* p : sc->ws) hipFree(p);
    sc->ws.clear();
    sc->cap_paths = sc->cap_pix = 0;
    sc->film = nullptr;
}

template <class T>
int ws_alloc(sortgpu_scene* sc, T** out, size_t count) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return hip_fail(sc->accel->ctx, e, "hipMalloc(path workspace)");
    sc->ws.push_back(p);
    *out = static_cast<T*>(p);
    return SORTGPU_OK;
}

int ensure_ws(sortgpu_scene* sc, uint32_t paths, uint32_t pix) {
    if (paths <= sc->cap_paths && pix <= sc->cap_pix) return SORTGPU_OK;
    free_ws(sc);
    paths = std::max(paths, 1u);
    pix = std::max(pix, 1u);
    int rc;
    PathState& ps = sc->ps;
    Queues& q = sc->q;
    if ((rc = ws_alloc(sc, &ps.T, paths)) || (rc = ws_alloc(sc, &ps.L, paths)) || (rc = ws_alloc(sc, &ps.c0, paths)) ||
        (rc = ws_alloc(sc, &ps.c1, paths)) || (rc = ws_alloc(sc, &ps.key, paths)) ||
        (rc = ws_alloc(sc, &ps.smp, paths)) || (rc = ws_alloc(sc, &ps.req, paths)) ||
        (rc = ws_alloc(sc, &q.ray[0], 2 * (size_t)paths)) || (rc = ws_alloc(sc, &q.ray[1], 2 * (size_t)paths)) ||
        (rc = ws_alloc(sc, &q.pid[0], paths)) || (rc = ws_alloc(sc, &q.pid[1], paths)) ||
        (rc = ws_alloc(sc, &q.hit, paths)) || (rc = ws_alloc(sc, &q.sray, 4 * (size_t)paths)) ||
        (rc = ws_alloc(sc, &q.ssid, 2 * (size_t)paths)) || (rc = ws_alloc(sc, &q.socc, 2 * (size_t)paths)) ||
        (rc = ws_alloc(sc, &q.count, 16)) || (rc = ws_alloc(sc, &q.totals, 4)) ||
        (rc = ws_alloc(sc, &sc->film, pix))) {
        free_ws(sc);
        return rc;
    }
    sc->cap_paths = paths;
    sc->cap_pix = pix;
    return SORTGPU_OK;
}

float lum(const float* c) { return 0.212671f * c[0] + 0.715160f * c[1] + 0.072169f * c[2]; }

PtLaunch make_launch(sortgpu_scene* sc, uint32_t max_depth, uint64_t seed) {
    PtLaunch L;
    std::memset(&L, 0, sizeof(L));
    sortgpu_accel* acc = sc->accel;
    L.trace.tree = acc->view;
    const uint32_t slot = acc->slot.fetch_add(1) % kCounterSlots;
    L.trace.counter = acc->d_counters + slot * kCounterStride;
    L.trace.error = acc->d_error;
    L.trace.stats = nullptr;
    apply_tuning(acc->ctx, L.trace);
    L.scenes = sc->d_scene;
    L.ps = sc->ps;
    L.q = sc->q;
    L.seed = (uint32_t)(seed ^ (seed >> 32));
    L.max_depth = max_depth;
    L.num_cus = acc->ctx->num_cus;
    return L;
}

// Paths in flight per batch: about 2M, whole samples of the frame at a time.
uint32_t samples_per_batch(uint32_t npix, uint32_t spp) {
    const uint64_t target = 1u << 21;
    uint64_t s = target / std::max<uint32_t>(npix, 1);
    s = std::max<uint64_t>(1, std::min<uint64_t>(s, spp));
    return (uint32_t)s;
}

}  // namespace

extern "C" {

int sortgpu_scene_create(sortgpu_accel* acc, const sortgpu_scene_desc* desc, sortgpu_scene** out) {
    if (!acc || !desc || !out) {
        set_error(acc ? acc->ctx : nullptr, "sortgpu_scene_create: null argument");
        return SORTGPU_ERR_ARG;
    }
    *out = nullptr;
    sortgpu_ctx* ctx = acc->ctx;
    if (ctx->host_only) {
        set_error(ctx, "sortgpu_scene_create: host-only context");
        return SORTGPU_ERR_STATE;
    }
    const uint32_t ntris = (uint32_t)(acc->idx.size() / 3), nverts = (uint32_t)(acc->pos.size() / 3);
    const uint32_t nquads = (uint32_t)acc->quads.size();
    if (desc->nmaterials && !desc->materials) {
        set_error(ctx, "sortgpu_scene_create: materials is NULL");
        return SORTGPU_ERR_ARG;
    }
    if (nquads && !desc->quad_radiance) {
        set_error(ctx, "sortgpu_scene_create: quad_radiance is NULL but the accelerator has quads");
        return SORTGPU_ERR_ARG;
    }
    if (desc->tri_material) {
        for (uint32_t i = 0; i < ntris; i++)
            if (desc->tri_material[i] >= std::max(desc->nmaterials, 1u)) {
                set_error(ctx, "sortgpu_scene_create: triangle %u has material %u (of %u)", i, desc->tri_material[i],
                          desc->nmaterials);
                return SORTGPU_ERR_ARG;
            }
    } else if (ntris && desc->nmaterials == 0) {
        set_error(ctx, "sortgpu_scene_create: triangles need at least one material");
        return SORTGPU_ERR_ARG;
    }
    for (uint32_t m = 0; m < desc->nmaterials; m++)
        if (desc->materials[m].type > SORTGPU_MAT_DISNEY) {
            set_error(ctx, "sortgpu_scene_create: material %u has unknown type %u", m, desc->materials[m].type);
            return SORTGPU_ERR_ARG;
        }
    int rc = ctx_bind(ctx);
    if (rc) return rc;
    sortgpu_scene* sc = new (std::nothrow) sortgpu_scene();
    if (!sc) return SORTGPU_ERR_OOM;
    sc->accel = acc;

    // materials + the default white Lambert
    std::vector<MatRec> mats(desc->nmaterials + 1);
    std::memset(mats.data(), 0, mats.size() * sizeof(MatRec));
    for (uint32_t m = 0; m < desc->nmaterials; m++) {
        const sortgpu_material& s = desc->materials[m];
        MatRec& d = mats[m];
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
    mats.back().type = MAT_LAMBERT;
    mats.back().base = v3(1.f, 1.f, 1.f);

    // lights: area quads in quad order, then the environment (Scene::genLightDistribution)
    std::vector<LightRec> lights;
    std::vector<float> power;
    std::vector<int32_t> quad_light(nquads);
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
        quad_light[qi] = (int32_t)lights.size();
        lights.push_back(l);
        power.push_back(l.area * lum(Lq) * 6.28318530717958647692f);  // AreaLight::Power
    }
    int32_t env_idx = -1;
    const float* E = desc->env_radiance;
    if (E[0] != 0.0f || E[1] != 0.0f || E[2] != 0.0f) {
        LightRec l;
        std::memset(&l, 0, sizeof(l));
        l.L = v3(E[0], E[1], E[2]);
        l.kind = 1;
        env_idx = (int32_t)lights.size();
        lights.push_back(l);
        const float* b = acc->host.bbox;
        const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
        const float r = std::sqrt(dx * dx + dy * dy + dz * dz) * 0.5f;
        power.push_back(r * r * 3.14159265358979323846f * lum(E));  // SkyLight::Power
    }
    const uint32_t nl = (uint32_t)lights.size();
    std::vector<float> cdf(nl + 1, 0.0f);  // Distribution1D (core/samplemethod.h:132-152)
    for (uint32_t i = 0; i < nl; i++) cdf[i + 1] = cdf[i] + power[i];
    const float sum = nl ? cdf[nl] : 0.0f;
    for (uint32_t i = 0; i <= nl && nl; i++) cdf[i] = sum != 0.0f ? cdf[i] / sum : (float)i / (float)nl;

    SceneDev& h = sc->host;
    std::memset(&h, 0, sizeof(h));
    h.tree = acc->view;
    h.ntris = ntris;
    h.nquads = nquads;
    h.nlights = nl;
    h.nmats = desc->nmaterials;
    h.env_light = env_idx;
    h.env_L = v3(E[0], E[1], E[2]);
    if ((rc = upload(sc, acc->idx.data(), acc->idx.size(), &h.idx, "upload(idx)")) ||
        (rc = upload(sc, acc->pos.data(), acc->pos.size(), &h.pos, "upload(pos)")) ||
        (rc = upload(sc, desc->normals, desc->normals ? 3 * (size_t)nverts : 0, &h.nrm, "upload(normals)")) ||
        (rc = upload(sc, desc->tangents, desc->tangents ? 3 * (size_t)nverts : 0, &h.tan, "upload(tangents)")) ||
        (rc = upload(sc, desc->tri_material, desc->tri_material ? (size_t)ntris : 0, &h.tri_mat, "upload(materials)")) ||
        (rc = upload(sc, mats.data(), mats.size(), &h.mats, "upload(material table)")) ||
        (rc = upload(sc, lights.data(), lights.size(), &h.lights, "upload(lights)")) ||
        (rc = upload(sc, cdf.data(), cdf.size(), &h.light_cdf, "upload(light cdf)")) ||
        (rc = upload(sc, quad_light.data(), quad_light.size(), &h.quad_light, "upload(quad lights)"))) {
        sortgpu_scene_destroy(sc);
        return rc;
    }
    const SceneDev* dsc = nullptr;
    if ((rc = upload(sc, &sc->host, 1, &dsc, "upload(scene)"))) {
        sortgpu_scene_destroy(sc);
        return rc;
    }
    sc->d_scene = const_cast<SceneDev*>(dsc);
    if (hipEventCreate(&sc->ev0) != hipSuccess || hipEventCreate(&sc->ev1) != hipSuccess) {
        set_error(ctx, "sortgpu_scene_create: hipEventCreate failed");
        sortgpu_scene_destroy(sc);
        return SORTGPU_ERR_HIP;
    }
    *out = sc;
    return SORTGPU_OK;
}

void sortgpu_scene_destroy(sortgpu_scene* sc) {
    if (!sc) return;
    ctx_bind(sc->accel->ctx);
    free_ws(sc);
    for (void* p : sc->allocs) hipFree(p);
    if (sc->ev0) hipEventDestroy(sc->ev0);
    if (sc->ev1) hipEventDestroy(sc->ev1);
    delete sc;
}

int sortgpu_render(sortgpu_scene* sc, const sortgpu_camera* cam, uint32_t width, uint32_t height, uint32_t spp,
                   uint32_t max_depth, uint64_t seed, float* d_film_rgb, void* stream) {
    if (!sc || !cam || !d_film_rgb || width == 0 || height == 0 || spp == 0) {
        set_error(sc ? sc->accel->ctx : nullptr, "sortgpu_render: bad argument");
        return SORTGPU_ERR_ARG;
    }
    if ((uint64_t)width * height > (1ull << 28)) {
        set_error(sc->accel->ctx, "sortgpu_render: image too large");
        return SORTGPU_ERR_ARG;
    }
    sortgpu_ctx* ctx = sc->accel->ctx;
    int rc = ctx_bind(ctx);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t npix = width * height;
    const uint32_t spb = samples_per_batch(npix, spp);
    if ((rc = ensure_ws(sc, npix * spb, npix))) return rc;

    // camera basis (ViewLookat + Perspective, camera/perspective.cpp:26-64; math/transform.h)
    CamDev cd;
    const V3 eye = v3(cam->eye[0], cam->eye[1], cam->eye[2]);
    V3 f = v3(cam->target[0] - eye.x, cam->target[1] - eye.y, cam->target[2] - eye.z);
    const float fl = std::sqrt(f.x * f.x + f.y * f.y + f.z * f.z);
    if (!(fl > 0.0f) || !(cam->fov_y > 0.0f) || !(cam->fov_y < 3.14159f)) {
        set_error(ctx, "sortgpu_render: degenerate camera");
        return SORTGPU_ERR_ARG;
    }
    f = f / fl;
    const V3 up = v3(cam->up[0], cam->up[1], cam->up[2]);
    V3 r = v3(up.y * f.z - up.z * f.y, up.z * f.x - up.x * f.z, up.x * f.y - up.y * f.x);  // cross(up, dir)
    const float rl = std::sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
    if (!(rl > 0.0f)) {
        set_error(ctx, "sortgpu_render: up is parallel to the view direction");
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

    PtLaunch L = make_launch(sc, max_depth, seed);
    hipError_t e = hipEventRecord(sc->ev0, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventRecord");
    if ((e = pt_zero_f4(sc->film, npix, s)) != hipSuccess) return hip_fail(ctx, e, "zero film");
    if ((e = pt_zero_f4(reinterpret_cast<float4*>(sc->q.totals), 1, s)) != hipSuccess) return hip_fail(ctx, e, "zero totals");
    for (uint32_t s0 = 0; s0 < spp; s0 += spb) {
        const uint32_t ns = std::min(spb, spp - s0);
        if ((e = pt_render_batch(L, cd, s0, ns, sc->film, s)) != hipSuccess) return hip_fail(ctx, e, "render batch");
    }
    if ((e = pt_resolve_film(sc->film, npix, d_film_rgb, s)) != hipSuccess) return hip_fail(ctx, e, "resolve film");
    if ((e = hipEventRecord(sc->ev1, s)) != hipSuccess) return hip_fail(ctx, e, "hipEventRecord");
    sc->last_paths = (uint64_t)npix * spp;
    return SORTGPU_OK;
}

int sortgpu_li(sortgpu_scene* sc, const sortgpu_ray* d_rays, const uint32_t* d_keys, uint32_t n, uint32_t max_depth,
               uint64_t seed, float* d_rgb, void* stream) {
    if (!sc) {
        set_error(nullptr, "sortgpu_li: scene is NULL");
        return SORTGPU_ERR_ARG;
    }
    sortgpu_ctx* ctx = sc->accel->ctx;
    if (n == 0) return SORTGPU_OK;
    if (!d_rays || !d_keys || !d_rgb) {
        set_error(ctx, "sortgpu_li: null pointer");
        return SORTGPU_ERR_ARG;
    }
    int rc = ctx_bind(ctx);
    if (rc) return rc;
    if ((rc = ensure_ws(sc, n, 1))) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    PtLaunch L = make_launch(sc, max_depth, seed);
    hipError_t e = hipEventRecord(sc->ev0, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventRecord");
    if ((e = pt_zero_f4(reinterpret_cast<float4*>(sc->q.totals), 1, s)) != hipSuccess) return hip_fail(ctx, e, "zero totals");
    if ((e = pt_li_rays(L, reinterpret_cast<const float4*>(d_rays), d_keys, n, d_rgb, s)) != hipSuccess)
        return hip_fail(ctx, e, "li");
    if ((e = hipEventRecord(sc->ev1, s)) != hipSuccess) return hip_fail(ctx, e, "hipEventRecord");
    sc->last_paths = n;
    return SORTGPU_OK;
}

int sortgpu_render_stats(sortgpu_scene* sc, uint64_t out[4]) {
    if (!sc || !out) return SORTGPU_ERR_ARG;
    sortgpu_ctx* ctx = sc->accel->ctx;
    int rc = ctx_bind(ctx);
    if (rc) return rc;
    hipError_t e = hipEventSynchronize(sc->ev1);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventSynchronize");
    unsigned long long t[2] = {0, 0};
    if (sc->q.totals) {
        e = hipMemcpy(t, sc->q.totals, sizeof(t), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpy(totals)");
    }
    float ms = 0.0f;
    e = hipEventElapsedTime(&ms, sc->ev0, sc->ev1);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventElapsedTime");
    uint32_t err = 0;
    e = hipMemcpy(&err, sc->accel->d_error, sizeof(err), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpy(error)");
    out[0] = t[0];
    out[1] = t[1];
    out[2] = sc->last_paths;
    out[3] = (uint64_t)(ms * 1000.0f);
    if (err) {
        set_error(ctx, "render: traversal stack overflow");
        return SORTGPU_ERR_STATE;
    }
    return SORTGPU_OK;
}

}  // extern "C"
