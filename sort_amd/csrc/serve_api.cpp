// C ABI of the render server (sortgpu_server_*).
#include <algorithm>
#include <new>
#include <string>

#include "../../include/sortgpu.h"
#include "api_internal.h"
#include "serve_device.h"
#include "serve_engine.h"

using namespace sortgpu;

struct sortgpu_server {
    sortgpu_ctx* ctx = nullptr;  // held (reference counted) until the server is destroyed
    std::unique_ptr<serve::Engine> eng;
    ~sortgpu_server() {
        eng.reset();
        ctx_release(ctx);
    }
};

namespace {
constexpr uint32_t kMaxServerScenes = 256;

int fail(sortgpu_ctx* ctx, int rc, const char* fn, const std::string& err) {
    set_error(ctx, "%s: %s", fn, err.c_str());
    return rc;
}
}  // namespace

extern "C" {

int sortgpu_server_create(sortgpu_ctx* ctx, const sortgpu_server_opts* opts, sortgpu_server** out) {
    if (!ctx || !out) {
        set_error(ctx, "sortgpu_server_create: null argument");
        return SORTGPU_ERR_ARG;
    }
    *out = nullptr;
    sortgpu_server_opts o{};
    if (opts) o = *opts;
    serve::Options eo;
    eo.budget = o.hbm_budget ? o.hbm_budget : (96ull << 30);
    eo.max_active = o.max_active ? o.max_active : 32;
    eo.inflight = o.inflight_waves ? o.inflight_waves : 3;
    eo.manual = o.manual != 0;
    eo.fail_alloc_at = o.fail_alloc_at;
    const uint32_t wave_paths = o.wave_paths ? o.wave_paths : (1u << 25);
    if (eo.max_active > 32 || eo.inflight > 4 || wave_paths > (1u << 26)) {
        set_error(ctx, "sortgpu_server_create: max_active 1..32, inflight_waves 1..4, wave_paths <= 2^26");
        return SORTGPU_ERR_ARG;
    }
    std::unique_ptr<serve::Device> dev;
    std::string err;
    if (o.host_device) {
        dev = serve::make_host_device(ctx, wave_paths, 1);
    } else {
        if (ctx->host_only) {
            set_error(ctx, "sortgpu_server_create: host-only context (use host_device for the stand-in)");
            return SORTGPU_ERR_STATE;
        }
        int rc = serve::make_hip_device(ctx, wave_paths, kMaxServerScenes, eo.inflight, o.graphs != 0, dev, err);
        if (rc) return fail(ctx, rc, "sortgpu_server_create", err);
    }
    sortgpu_server* s = new (std::nothrow) sortgpu_server();
    if (!s) return SORTGPU_ERR_OOM;
    s->ctx = ctx;
    ctx_acquire(ctx);
    s->eng.reset(new (std::nothrow) serve::Engine(std::move(dev), eo));
    if (!s->eng) {
        delete s;
        return SORTGPU_ERR_OOM;
    }
    int rc = s->eng->start(err);
    if (rc) {
        delete s;
        return fail(ctx, rc, "sortgpu_server_create", err);
    }
    *out = s;
    return SORTGPU_OK;
}

void sortgpu_server_destroy(sortgpu_server* srv) { delete srv; }

int sortgpu_server_add_scene(sortgpu_server* srv, const sortgpu_accel* accel, const sortgpu_scene_desc* desc,
                             uint32_t* scene_id) {
    if (!srv || !accel || !desc || !scene_id) {
        set_error(srv ? srv->ctx : nullptr, "sortgpu_server_add_scene: null argument");
        return SORTGPU_ERR_ARG;
    }
    std::string err;
    int rc = srv->eng->add_scene(accel, desc, scene_id, err);
    return rc ? fail(srv->ctx, rc, "sortgpu_server_add_scene", err) : SORTGPU_OK;
}

int sortgpu_server_submit(sortgpu_server* srv, const sortgpu_request* req, uint64_t* id) {
    if (!srv || !req || !id) {
        set_error(srv ? srv->ctx : nullptr, "sortgpu_server_submit: null argument");
        return SORTGPU_ERR_ARG;
    }
    std::string err;
    int rc = srv->eng->submit(*req, id, err);
    return rc ? fail(srv->ctx, rc, "sortgpu_server_submit", err) : SORTGPU_OK;
}

int sortgpu_server_cancel(sortgpu_server* srv, uint64_t id) {
    if (!srv) return SORTGPU_ERR_ARG;
    int rc = srv->eng->cancel(id);
    return rc ? fail(srv->ctx, rc, "sortgpu_server_cancel", "unknown request") : SORTGPU_OK;
}

int sortgpu_server_info(sortgpu_server* srv, uint64_t id, sortgpu_request_info* out) {
    if (!srv || !out) return SORTGPU_ERR_ARG;
    int rc = srv->eng->info(id, out);
    return rc ? fail(srv->ctx, rc, "sortgpu_server_info", "unknown request") : SORTGPU_OK;
}

int sortgpu_server_wait(sortgpu_server* srv, uint64_t id, double timeout_s, sortgpu_request_info* out) {
    if (!srv) return SORTGPU_ERR_ARG;
    int rc = srv->eng->wait(id, timeout_s, out);
    return rc ? fail(srv->ctx, rc, "sortgpu_server_wait", "unknown request") : SORTGPU_OK;
}

int sortgpu_server_result(sortgpu_server* srv, uint64_t id, float* rgb, uint64_t n_floats) {
    if (!srv) return SORTGPU_ERR_ARG;
    std::string err;
    int rc = srv->eng->result(id, rgb, n_floats, err);
    return rc ? fail(srv->ctx, rc, "sortgpu_server_result", err) : SORTGPU_OK;
}

int sortgpu_server_message(sortgpu_server* srv, uint64_t id, char* buf, uint32_t cap) {
    if (!srv || !buf || cap == 0) return SORTGPU_ERR_ARG;
    std::string msg;
    int rc = srv->eng->message(id, msg);
    if (rc) return fail(srv->ctx, rc, "sortgpu_server_message", "unknown request");
    const size_t n = std::min<size_t>(msg.size(), cap - 1);
    msg.copy(buf, n);
    buf[n] = 0;
    return SORTGPU_OK;
}

int sortgpu_server_release(sortgpu_server* srv, uint64_t id) {
    if (!srv) return SORTGPU_ERR_ARG;
    int rc = srv->eng->release(id);
    return rc ? fail(srv->ctx, rc, "sortgpu_server_release", "unknown request") : SORTGPU_OK;
}

int sortgpu_server_step(sortgpu_server* srv, int* busy) {
    if (!srv) return SORTGPU_ERR_ARG;
    std::string err;
    int rc = srv->eng->step(busy, err);
    return rc ? fail(srv->ctx, rc, "sortgpu_server_step", err) : SORTGPU_OK;
}

int sortgpu_server_drain(sortgpu_server* srv, double timeout_s) {
    if (!srv) return SORTGPU_ERR_ARG;
    int rc = srv->eng->drain(timeout_s);
    return rc ? fail(srv->ctx, rc, "sortgpu_server_drain", "timed out") : SORTGPU_OK;
}

int sortgpu_server_stats(sortgpu_server* srv, uint64_t out[16]) {
    if (!srv || !out) return SORTGPU_ERR_ARG;
    srv->eng->stats(out);
    return SORTGPU_OK;
}

}  // extern "C"
