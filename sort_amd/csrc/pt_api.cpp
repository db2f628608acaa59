// C ABI of the path tracer: scene upload, render, batched Li, statistics. One scene object owns
// its shading data, a path workspace and a film; calls on one scene object are not thread-safe.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/sortgpu.h"
#include "pt_host.h"

using namespace sortgpu;

struct sortgpu_scene {
    sortgpu_accel* accel = nullptr;
    sortgpu_ctx* ctx = nullptr;   // held (reference counted) until the scene is destroyed
    SceneDev host{};              // pointers are device addresses
    SceneDev* d_scene = nullptr;  // one-entry scene table
    DeviceAllocs mem;             // shading data + scene table
    // A frame's waves take turns on three internal streams, each with its own workspace, so the
    // latency-bound last bounces of some waves overlap the first bounces of others; the caller's
    // stream waits for all of them at the end, and film updates run in wave order (events).
    static constexpr int kStreams = 3;
    PathWorkspace ws[kStreams];
    hipStream_t rs[kStreams] = {};
    hipEvent_t ev_fork = nullptr, ev_film = nullptr, ev_join[kStreams] = {};
    DeviceAllocs film_mem;
    float4* film = nullptr;
    uint32_t cap_pix = 0;
    ReqRing ring;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t last_paths = 0;
    // stage timing: six events per bounce (sortgpu_render_set_timing); stage_used = events the
    // last render recorded
    static constexpr uint32_t kStageBounces = 1024;
    std::vector<hipEvent_t> stage_ev;
    bool timing = false;
    uint32_t stage_used = 0;
};

namespace {

int ensure_film(sortgpu_scene* sc, uint32_t npix) {
    if (npix <= sc->cap_pix) return SORTGPU_OK;
    sc->film_mem.free_all();
    sc->film = nullptr;
    sc->cap_pix = 0;
    void* p = nullptr;
    int rc = sc->film_mem.alloc(sc->accel->ctx, &p, (size_t)npix * sizeof(float4), "hipMalloc(film)");
    if (rc) return rc;
    sc->film = static_cast<float4*>(p);
    sc->cap_pix = npix;
    return SORTGPU_OK;
}

// The internal streams and their events, made on first use.
int ensure_streams(sortgpu_scene* sc) {
    if (sc->rs[0]) return SORTGPU_OK;
    hipError_t e = hipSuccess;
    for (int k = 0; k < sortgpu_scene::kStreams && e == hipSuccess; k++) {
        e = hipStreamCreateWithFlags(&sc->rs[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sc->ev_join[k], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sc->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sc->ev_film, hipEventDisableTiming);
    return e == hipSuccess ? SORTGPU_OK : hip_fail(sc->ctx, e, "render streams");
}

PtLaunch make_launch(sortgpu_scene* sc, const WaveDesc* d_wave, uint32_t max_depth, int k = 0) {
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
    L.wave = d_wave;
    L.trees = nullptr;
    L.ps = sc->ws[k].ps;
    L.q = sc->ws[k].q;
    L.max_depth = max_depth;
    L.num_cus = acc->ctx->num_cus;
    L.transp = sc->host.transp;
    L.sort_min = shade_sort_min(acc->ctx->tuning);
    return L;
}

// Paths in flight per wave: about 8M, whole samples of the frame at a time (a frame of 8M
// pixels or more runs one sample per wave), and at least one wave per internal stream as long as
// each wave keeps 2^18 paths or more (a 1M-path frame runs as three waves whose thin last bounces
// overlap). With three waves in flight, the workspace of a scene object is at most 3 x 248 B per
// path slot: 6.2 GB for waves of 8M paths.
uint32_t samples_per_wave(uint32_t npix, uint32_t spp) {
    const uint64_t target = 1u << 23, min_paths = 1u << 18;
    const uint64_t np = std::max<uint32_t>(npix, 1);
    uint64_t s = target / np;
    s = std::min<uint64_t>(s, (spp + sortgpu_scene::kStreams - 1) / sortgpu_scene::kStreams);
    s = std::max<uint64_t>(s, (min_paths + np - 1) / np);
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
    int rc = ctx_bind(ctx);
    if (rc) return rc;
    sortgpu_scene* sc = new (std::nothrow) sortgpu_scene();
    if (!sc) return SORTGPU_ERR_OOM;
    sc->accel = acc;
    sc->ctx = ctx;
    ctx_acquire(ctx);
    SceneImage im;
    void* block = nullptr;
    void* dsc = nullptr;
    if ((rc = prepare_scene(ctx, acc, desc, false, im)) ||
        (rc = sc->mem.alloc(ctx, &block, im.block_bytes(), "hipMalloc(scene)")) ||
        (rc = im.place(ctx, block, sc->host, nullptr))) {
        sortgpu_scene_destroy(sc);
        return rc;
    }
    if (hipStreamSynchronize(nullptr) != hipSuccess) {  // `im` goes out of scope
        set_error(ctx, "sortgpu_scene_create: scene upload failed");
        sortgpu_scene_destroy(sc);
        return SORTGPU_ERR_HIP;
    }
    sc->host.tree = acc->view;
    if ((rc = sc->mem.upload(ctx, &sc->host, sizeof(SceneDev), &dsc, "upload(scene)")) || (rc = sc->ring.init(ctx))) {
        sortgpu_scene_destroy(sc);
        return rc;
    }
    sc->d_scene = static_cast<SceneDev*>(dsc);
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
    sortgpu_ctx* ctx = sc->ctx;
    ctx_bind(ctx);
    sc->ring.release();
    if (sc->ev1) hipEventSynchronize(sc->ev1);
    for (int k = 0; k < sortgpu_scene::kStreams; k++) {
        if (sc->rs[k]) {
            hipStreamSynchronize(sc->rs[k]);
            hipStreamDestroy(sc->rs[k]);
        }
        if (sc->ev_join[k]) hipEventDestroy(sc->ev_join[k]);
        sc->ws[k].release();
    }
    if (sc->ev_fork) hipEventDestroy(sc->ev_fork);
    if (sc->ev_film) hipEventDestroy(sc->ev_film);
    sc->film_mem.free_all();
    sc->mem.free_all();
    if (sc->ev0) hipEventDestroy(sc->ev0);
    if (sc->ev1) hipEventDestroy(sc->ev1);
    for (hipEvent_t ev : sc->stage_ev) hipEventDestroy(ev);
    delete sc;
    ctx_release(ctx);
}

}  // extern "C"

namespace {

// The frame loop of sortgpu_render; with a dump, also records the extension rays and hits.
int render_impl(sortgpu_scene* sc, const sortgpu_camera* cam, uint32_t width, uint32_t height, uint32_t spp,
                uint32_t max_depth, uint64_t seed, float* d_film_rgb, const RayDumpDev* dump, void* stream) {
    sortgpu_ctx* ctx = sc->accel->ctx;
    int rc = ctx_bind(ctx);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t npix = width * height;
    const uint32_t spw = samples_per_wave(npix, spp);
    WaveDesc wd;
    std::memset(&wd, 0, sizeof(wd));
    ReqDev& R = wd.reqs[0];
    if ((rc = make_camera(ctx, cam, width, height, R.cam))) return rc;
    const uint32_t waves = (spp + spw - 1) / spw;
    const int nstreams = waves > 1 ? sortgpu_scene::kStreams : 1;
    for (int k = 0; k < nstreams; k++)
        if ((rc = sc->ws[k].ensure(ctx, npix * spw))) return rc;
    if ((rc = ensure_film(sc, npix)) || (nstreams > 1 && (rc = ensure_streams(sc)))) return rc;
    R.scene = 0;
    R.seed = fold_seed(seed);
    R.max_depth = max_depth;
    R.npix = npix;
    R.film = sc->film;
    R.clamp = sc->host.clamp;
    hipError_t e = hipEventRecord(sc->ev0, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventRecord");
    sc->stage_used = 0;
    if ((e = pt_zero_f4(sc->film, npix, s)) != hipSuccess) return hip_fail(ctx, e, "zero film");
    for (int k = 0; k < sortgpu_scene::kStreams; k++)
        if (sc->ws[k].q.totals &&
            (e = pt_zero_f4(reinterpret_cast<float4*>(sc->ws[k].q.totals), 1, s)) != hipSuccess)
            return hip_fail(ctx, e, "zero totals");
    if (nstreams > 1) {  // fork: the internal streams start after the caller's work so far
        if ((e = hipEventRecord(sc->ev_fork, s)) != hipSuccess) return hip_fail(ctx, e, "hipEventRecord");
        for (int k = 0; k < nstreams; k++)
            if ((e = hipStreamWaitEvent(sc->rs[k], sc->ev_fork, 0)) != hipSuccess)
                return hip_fail(ctx, e, "hipStreamWaitEvent");
    }
    uint32_t w = 0;
    for (uint32_t s0 = 0; s0 < spp; s0 += spw, w++) {
        const uint32_t ns = std::min(spw, spp - s0);
        const int k = (int)(w % (uint32_t)nstreams);
        hipStream_t st = nstreams > 1 ? sc->rs[k] : s;
        wd.ct.n = 1;
        wd.ct.total_paths = npix * ns;
        wd.ct.total_pix = npix;
        wd.ct.c[0] = ChunkDev{0u, s0, ns, npix, 0u, 0u};
        const WaveDesc* d_wave = nullptr;
        if ((rc = sc->ring.push(ctx, wd, st, nullptr, &d_wave))) return rc;
        PtLaunch L = make_launch(sc, d_wave, max_depth, k);
        L.paths_cap = npix * ns;
        L.pix_cap = npix;
        if (dump) L.dump = *dump;
        if (sc->timing) {
            L.stage_ev = sc->stage_ev.data();
            L.stage_cap = (uint32_t)sc->stage_ev.size();
            L.stage_used = &sc->stage_used;
        }
        if ((e = pt_wave_paths(L, st)) != hipSuccess) return hip_fail(ctx, e, "render wave");
        // film updates in wave order, so the film is summed in sample order
        if (w > 0 && nstreams > 1 && (e = hipStreamWaitEvent(st, sc->ev_film, 0)) != hipSuccess)
            return hip_fail(ctx, e, "hipStreamWaitEvent");
        if ((e = pt_wave_film(L, st)) != hipSuccess) return hip_fail(ctx, e, "render film");
        if (nstreams > 1 && (e = hipEventRecord(sc->ev_film, st)) != hipSuccess)
            return hip_fail(ctx, e, "hipEventRecord");
        if ((rc = sc->ring.done(ctx, st))) return rc;
    }
    if (nstreams > 1) {  // join: the caller's stream waits for both internal streams
        for (int k = 0; k < nstreams; k++)
            if ((e = hipEventRecord(sc->ev_join[k], sc->rs[k])) != hipSuccess ||
                (e = hipStreamWaitEvent(s, sc->ev_join[k], 0)) != hipSuccess)
                return hip_fail(ctx, e, "stream join");
    }
    if (d_film_rgb && (e = pt_resolve_film(sc->film, npix, d_film_rgb, s)) != hipSuccess)
        return hip_fail(ctx, e, "resolve film");
    if ((e = hipEventRecord(sc->ev1, s)) != hipSuccess) return hip_fail(ctx, e, "hipEventRecord");
    sc->last_paths = (uint64_t)npix * spp;
    return SORTGPU_OK;
}

int check_frame(sortgpu_scene* sc, const sortgpu_camera* cam, uint32_t width, uint32_t height, uint32_t spp,
                const char* fn) {
    if (!sc || !cam || width == 0 || height == 0 || spp == 0) {
        set_error(sc ? sc->accel->ctx : nullptr, "%s: bad argument", fn);
        return SORTGPU_ERR_ARG;
    }
    if ((uint64_t)width * height > (1ull << 24)) {
        set_error(sc->accel->ctx, "%s: image too large (at most 2^24 pixels)", fn);
        return SORTGPU_ERR_ARG;
    }
    return SORTGPU_OK;
}

}  // namespace

extern "C" {

int sortgpu_render(sortgpu_scene* sc, const sortgpu_camera* cam, uint32_t width, uint32_t height, uint32_t spp,
                   uint32_t max_depth, uint64_t seed, float* d_film_rgb, void* stream) {
    int rc = check_frame(sc, cam, width, height, spp, "sortgpu_render");
    if (rc) return rc;
    if (!d_film_rgb) {
        set_error(sc->accel->ctx, "sortgpu_render: d_film_rgb is NULL");
        return SORTGPU_ERR_ARG;
    }
    return render_impl(sc, cam, width, height, spp, max_depth, seed, d_film_rgb, nullptr, stream);
}

int sortgpu_render_ray_dump(sortgpu_scene* sc, const sortgpu_camera* cam, uint32_t width, uint32_t height,
                            uint32_t spp, uint32_t max_depth, uint64_t seed, uint32_t pix_begin, uint32_t pix_end,
                            sortgpu_ray* d_rays, sortgpu_hit* d_hits, uint8_t* d_valid, void* stream) {
    int rc = check_frame(sc, cam, width, height, spp, "sortgpu_render_ray_dump");
    if (rc) return rc;
    sortgpu_ctx* ctx = sc->accel->ctx;
    if (!d_rays || !d_hits || !d_valid || pix_begin >= pix_end || pix_end > width * height || max_depth == 0 ||
        (uint64_t)(pix_end - pix_begin) * spp * max_depth > (1ull << 31)) {
        set_error(ctx, "sortgpu_render_ray_dump: null buffer, empty or out-of-frame pixel range, or over 2^31 slots");
        return SORTGPU_ERR_ARG;
    }
    if ((rc = ctx_bind(ctx))) return rc;
    const size_t n = (size_t)(pix_end - pix_begin) * spp * max_depth;
    hipError_t e = hipMemsetAsync(d_valid, 0, n, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(ctx, e, "clear dump flags");
    RayDumpDev d{reinterpret_cast<float4*>(d_rays), reinterpret_cast<float4*>(d_hits), d_valid, pix_begin, pix_end,
                 spp, max_depth};
    return render_impl(sc, cam, width, height, spp, max_depth, seed, nullptr, &d, stream);
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
    if ((rc = sc->ws[0].ensure(ctx, n))) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    WaveDesc wd;
    std::memset(&wd, 0, sizeof(wd));
    wd.ct.n = 1;
    ReqDev& R = wd.reqs[0];
    R.scene = 0;
    R.seed = fold_seed(seed);
    R.max_depth = max_depth;
    const WaveDesc* d_wave = nullptr;
    hipError_t e = hipEventRecord(sc->ev0, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventRecord");
    if ((rc = sc->ring.push(ctx, wd, s, nullptr, &d_wave))) return rc;
    PtLaunch L = make_launch(sc, d_wave, max_depth);
    for (int k = 0; k < sortgpu_scene::kStreams; k++)
        if (sc->ws[k].q.totals &&
            (e = pt_zero_f4(reinterpret_cast<float4*>(sc->ws[k].q.totals), 1, s)) != hipSuccess)
            return hip_fail(ctx, e, "zero totals");
    if ((e = pt_li_rays(L, reinterpret_cast<const float4*>(d_rays), d_keys, n, d_rgb, s)) != hipSuccess)
        return hip_fail(ctx, e, "li");
    if ((rc = sc->ring.done(ctx, s))) return rc;
    if ((e = hipEventRecord(sc->ev1, s)) != hipSuccess) return hip_fail(ctx, e, "hipEventRecord");
    sc->last_paths = n;
    return SORTGPU_OK;
}

int sortgpu_render_set_timing(sortgpu_scene* sc, int on) {
    if (!sc) return SORTGPU_ERR_ARG;
    sortgpu_ctx* ctx = sc->accel->ctx;
    int rc = ctx_bind(ctx);
    if (rc) return rc;
    if (on && sc->stage_ev.empty()) {
        std::vector<hipEvent_t> evs;
        try {
            evs.reserve(6 * (size_t)sortgpu_scene::kStageBounces);
        } catch (const std::bad_alloc&) {
            set_error(ctx, "render timing: host allocation failed");
            return SORTGPU_ERR_OOM;
        }
        for (uint32_t i = 0; i < 6 * sortgpu_scene::kStageBounces; i++) {
            hipEvent_t ev = nullptr;
            const hipError_t e = hipEventCreate(&ev);
            if (e != hipSuccess) {
                for (hipEvent_t x : evs) hipEventDestroy(x);
                return hip_fail(ctx, e, "hipEventCreate(stage timing)");
            }
            evs.push_back(ev);
        }
        sc->stage_ev.swap(evs);
    }
    sc->timing = on != 0;
    sc->stage_used = 0;
    return SORTGPU_OK;
}

int sortgpu_render_stage_times(sortgpu_scene* sc, float out_ms[3], uint32_t* bounces) {
    if (!sc || !out_ms) return SORTGPU_ERR_ARG;
    sortgpu_ctx* ctx = sc->accel->ctx;
    int rc = ctx_bind(ctx);
    if (rc) return rc;
    out_ms[0] = out_ms[1] = out_ms[2] = 0.0f;
    if (bounces) *bounces = 0;
    if (sc->stage_used == 0) return SORTGPU_OK;
    hipError_t e = hipEventSynchronize(sc->ev1);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventSynchronize");
    double acc[3] = {0.0, 0.0, 0.0};
    for (uint32_t b = 0; b + 6 <= sc->stage_used; b += 6) {
        for (int k = 0; k < 3; k++) {
            float ms = 0.0f;
            if ((e = hipEventElapsedTime(&ms, sc->stage_ev[b + 2 * k], sc->stage_ev[b + 2 * k + 1])) != hipSuccess)
                return hip_fail(ctx, e, "hipEventElapsedTime(stage)");
            acc[k] += ms;
        }
    }
    for (int k = 0; k < 3; k++) out_ms[k] = (float)acc[k];
    if (bounces) *bounces = sc->stage_used / 6;
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
    for (int k = 0; k < sortgpu_scene::kStreams; k++) {
        if (!sc->ws[k].q.totals) continue;
        unsigned long long tk[2] = {0, 0};
        e = hipMemcpy(tk, sc->ws[k].q.totals, sizeof(tk), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpy(totals)");
        t[0] += tk[0];
        t[1] += tk[1];
    }
    float ms = 0.0f;
    e = hipEventElapsedTime(&ms, sc->ev0, sc->ev1);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventElapsedTime");
    out[0] = t[0];
    out[1] = t[1];
    out[2] = sc->last_paths;
    out[3] = (uint64_t)(ms * 1000.0f);
    return report_flags(ctx, sc->accel->d_error, nullptr, "render", nullptr);
}

}  // extern "C"
