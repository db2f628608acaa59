// C ABI implementation: context, accelerator build/upload, ray queries.
// Declarations and reference citations: include/sortgpu.h.
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "../../include/sortgpu.h"
#include "api_internal.h"
#include "bvh_build.h"
#include "trace_launch.h"

using namespace sortgpu;

namespace {
thread_local std::string g_lib_error;
}

namespace sortgpu {

void set_error(sortgpu_ctx* ctx, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) {
        std::lock_guard<std::mutex> l(ctx->mu);
        ctx->err = buf;
    }
    g_lib_error = buf;
}

int hip_fail(sortgpu_ctx* ctx, hipError_t e, const char* what) {
    set_error(ctx, "%s: %s", what, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? SORTGPU_ERR_OOM : SORTGPU_ERR_HIP;
}

void apply_tuning(const sortgpu_ctx* ctx, TraceArgs& a) {
    a.refill = ctx->tuning.refill;
    a.lds_entries = ctx->tuning.lds_entries;
    a.blocks_per_cu = ctx->tuning.blocks_per_cu;
    a.leaf_batch = ctx->tuning.leaf_batch;
    a.leaf_min = ctx->tuning.leaf_min;
}

void ctx_acquire(sortgpu_ctx* ctx) {
    if (ctx) ctx->refs.fetch_add(1);
}

void ctx_release(sortgpu_ctx* ctx) {
    if (ctx && ctx->refs.fetch_sub(1) == 1) delete ctx;
}

int ctx_bind(sortgpu_ctx* ctx) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipGetDevice");
    if (cur != ctx->device) {
        e = hipSetDevice(ctx->device);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    }
    return SORTGPU_OK;
}

int report_flags(sortgpu_ctx* ctx, uint32_t* d_flags, hipStream_t s, const char* who, uint32_t* flags_out) {
    // one taker at a time: the staging word d_flags[1] is shared by every check of the object
    static std::mutex take_mu;
    uint32_t f = 0;
    hipError_t e;
    {
        std::lock_guard<std::mutex> l(take_mu);
        e = take_flags(d_flags, s, &f);
    }
    if (flags_out) *flags_out = f;
    if (e != hipSuccess) return hip_fail(ctx, e, "error flags");
    if (!f) return SORTGPU_OK;
    std::string msg;
    if (f & kFlagStackOverflow) msg += "; traversal stack overflow (those rays' results are incomplete)";
    if (f & kFlagBadScene) msg += "; a ray named a scene outside the pool (it missed)";
    if (f & kFlagAttenCap) msg += "; an attenuation walk passed 4096 surfaces and stopped there";
    if (f & ~(kFlagStackOverflow | kFlagBadScene | kFlagAttenCap)) msg += "; unknown flag";
    set_error(ctx, "%s: flags %u%s", who, f, msg.c_str());
    return SORTGPU_ERR_STATE;
}

}  // namespace sortgpu

extern "C" {

int sortgpu_ctx_create(int device, sortgpu_ctx** out) {
    if (!out) {
        set_error(nullptr, "sortgpu_ctx_create: out is NULL");
        return SORTGPU_ERR_ARG;
    }
    *out = nullptr;
    sortgpu_ctx* ctx = new (std::nothrow) sortgpu_ctx();
    if (!ctx) return SORTGPU_ERR_OOM;
    ctx->device = device;
    ctx->host_only = device < 0;
    if (!ctx->host_only) {
        int count = 0;
        hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess || device >= count) {
            set_error(nullptr, "sortgpu_ctx_create: device %d not available (%s)", device,
                      e != hipSuccess ? hipGetErrorString(e) : "index out of range");
            delete ctx;
            return SORTGPU_ERR_HIP;
        }
        e = hipSetDevice(device);
        if (e != hipSuccess) {
            int rc = hip_fail(nullptr, e, "hipSetDevice");
            delete ctx;
            return rc;
        }
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, device);
        if (e != hipSuccess) {
            int rc = hip_fail(nullptr, e, "hipGetDeviceProperties");
            delete ctx;
            return rc;
        }
        ctx->num_cus = prop.multiProcessorCount;
        ctx->hbm_bytes = prop.totalGlobalMem;
        std::snprintf(ctx->arch, sizeof(ctx->arch), "%s", prop.gcnArchName);
    }
    *out = ctx;
    return SORTGPU_OK;
}

// Drops the creator's reference; the context lives on while accelerators, pools, scenes or
// servers made from it do.
void sortgpu_ctx_destroy(sortgpu_ctx* ctx) { ctx_release(ctx); }

int sortgpu_ctx_set_tuning(sortgpu_ctx* ctx, const sortgpu_tuning* t) {
    if (!ctx || !t) {
        set_error(ctx, "sortgpu_ctx_set_tuning: null argument");
        return SORTGPU_ERR_ARG;
    }
    if (t->refill > 64 || (t->lds_entries != 0 && t->lds_entries != 8 && t->lds_entries != 16) || t->leaf_batch > 64 ||
        t->leaf_min > 64 || t->reorder > 2 || t->shade_order > 2) {
        set_error(ctx,
                  "sortgpu_ctx_set_tuning: refill, leaf_batch and leaf_min must be <= 64, lds_entries 0, 8 or 16, "
                  "reorder and shade_order 0, 1 or 2");
        return SORTGPU_ERR_ARG;
    }
    ctx->tuning = *t;
    return SORTGPU_OK;
}

const char* sortgpu_last_error(const sortgpu_ctx* ctx) {
    if (ctx) return ctx->err.c_str();
    return g_lib_error.c_str();
}

int sortgpu_accel_build(sortgpu_ctx* ctx, const float* pos, uint32_t nverts, const uint32_t* idx,
                        uint32_t ntris, const sortgpu_quad* quads, uint32_t nquads,
                        const sortgpu_build_opts* opts, sortgpu_accel** out) {
    if (!ctx || !out) {
        set_error(ctx, "sortgpu_accel_build: ctx and out must be non-NULL");
        return SORTGPU_ERR_ARG;
    }
    *out = nullptr;
    BuildInput in;
    in.pos = pos;
    in.nverts = nverts;
    in.idx = idx;
    in.ntris = ntris;
    static_assert(sizeof(sortgpu_quad) == sizeof(QuadIn), "quad record mismatch");
    in.quads = reinterpret_cast<const QuadIn*>(quads);
    in.nquads = nquads;
    if (opts) {
        if (opts->width) in.width = opts->width;
        if (opts->max_depth) in.max_depth = opts->max_depth;
        if (opts->max_leaf) in.max_leaf = opts->max_leaf;
        in.threads = opts->build_threads;
    }
    sortgpu_accel* acc = new (std::nothrow) sortgpu_accel();
    if (!acc) return SORTGPU_ERR_OOM;
    acc->ctx = ctx;
    std::string err;
    try {
        if (!build_tree(in, acc->host, err)) {
            set_error(ctx, "sortgpu_accel_build: %s", err.c_str());
            delete acc;
            return SORTGPU_ERR_ARG;
        }
    } catch (const std::bad_alloc&) {
        set_error(ctx, "sortgpu_accel_build: host allocation failed");
        delete acc;
        return SORTGPU_ERR_OOM;
    } catch (...) {
        set_error(ctx, "sortgpu_accel_build: unexpected failure");
        delete acc;
        return SORTGPU_ERR_ARG;
    }
    if (acc->host.max_stack > trace_max_stack_supported()) {
        set_error(ctx, "sortgpu_accel_build: tree depth %u needs %u stack entries (limit %u)", acc->host.depth,
                  acc->host.max_stack, trace_max_stack_supported());
        delete acc;
        return SORTGPU_ERR_ARG;
    }
    try {
        acc->pos.assign(pos, pos + 3 * (size_t)nverts);
        acc->idx.assign(idx, idx + 3 * (size_t)ntris);
        acc->quads.assign(quads, quads + nquads);
    } catch (const std::bad_alloc&) {
        set_error(ctx, "sortgpu_accel_build: host allocation failed");
        delete acc;
        return SORTGPU_ERR_OOM;
    }
    TreeView& v = acc->view;
    std::memset(&v, 0, sizeof(v));
    v.root = acc->host.root_ref;
    v.valid = acc->host.valid ? 1u : 0u;
    v.max_stack = acc->host.max_stack;
    v.width = acc->host.width;
    std::memcpy(v.bbox, acc->host.bbox, sizeof(v.bbox));
    if (!ctx->host_only) {
        int rc = upload_accel(acc);
        if (rc != SORTGPU_OK) {
            delete acc;
            return rc;
        }
    }
    ctx_acquire(ctx);
    *out = acc;
    return SORTGPU_OK;
}

void sortgpu_accel_destroy(sortgpu_accel* acc) {
    if (!acc) return;
    sortgpu_ctx* ctx = acc->ctx;
    free_accel_device(acc);
    delete acc;
    ctx_release(ctx);
}

int sortgpu_accel_valid(const sortgpu_accel* acc) { return acc && acc->host.valid ? 1 : 0; }

int sortgpu_accel_bbox(const sortgpu_accel* acc, float out6[6]) {
    if (!acc || !out6) return SORTGPU_ERR_ARG;
    std::memcpy(out6, acc->host.bbox, 6 * sizeof(float));
    return SORTGPU_OK;
}

int sortgpu_accel_info(const sortgpu_accel* acc, uint64_t out[8]) {
    if (!acc || !out) return SORTGPU_ERR_ARG;
    const BuiltTree& h = acc->host;
    out[0] = h.width;
    out[1] = h.depth;
    out[2] = h.n_inner;
    out[3] = h.n_leaves;
    out[4] = h.nprims;
    out[5] = h.max_stack;
    out[6] = h.node_bytes.size() + h.prims.size() * sizeof(PrimRec) + h.quads.size() * sizeof(QuadRec);
    out[7] = (uint64_t)(h.build_ms + 0.5);
    return SORTGPU_OK;
}

int sortgpu_accel_prim_order(const sortgpu_accel* acc, uint32_t* out, uint64_t cap) {
    if (!acc || (!out && cap)) return SORTGPU_ERR_ARG;
    const auto& po = acc->host.prim_order;
    const uint64_t n = po.size() < cap ? po.size() : cap;
    if (n) std::memcpy(out, po.data(), n * sizeof(uint32_t));
    return SORTGPU_OK;
}

int64_t sortgpu_accel_tree_dump(const sortgpu_accel* acc, uint32_t* out, uint64_t cap) {
    if (!acc) return SORTGPU_ERR_ARG;
    const auto& d = acc->host.dump;
    if (out) {
        const uint64_t n = d.size() < cap ? d.size() : cap;
        if (n) std::memcpy(out, d.data(), n * sizeof(uint32_t));
    }
    return (int64_t)d.size();
}

}  // extern "C"

// The context's pool for traversal-order buffers: created on first use, on the context's device,
// keeping freed memory for the next batch (release threshold at its maximum). Null if the device
// has no stream-ordered allocator.
hipMemPool_t sortgpu::order_pool(sortgpu_ctx* ctx) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (!ctx->order_pool) {
        hipMemPoolProps p{};
        p.allocType = hipMemAllocationTypePinned;
        p.location.type = hipMemLocationTypeDevice;
        p.location.id = ctx->device;
        hipMemPool_t pool = nullptr;
        if (hipMemPoolCreate(&pool, &p) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        ctx->order_pool = pool;
    }
    return ctx->order_pool;
}

extern "C" {

// Coherence order for a large batch on `s` (the tuning's reorder knob decides), or null for index
// order. An order that cannot be had (no pool, no memory) only costs speed, so it is not an error.
// Results never depend on it; `ro` is released on `s` after the launch that reads it.
static const uint32_t* traversal_order(sortgpu_accel* acc, const sortgpu_ray* d_rays, uint32_t n, hipStream_t s,
                                       RayOrder& ro) {
    sortgpu_ctx* ctx = acc->ctx;
    const uint32_t reorder = ctx->tuning.reorder;
    if (reorder == 1 || (reorder != 2 && n < kReorderMinRays) || n > 0x7fffffffu) return nullptr;
    hipMemPool_t pool = order_pool(ctx);
    if (pool && order_rays(pool, d_rays, n, acc->view.bbox, s, ro) == hipSuccess) return ro.order;
    (void)hipGetLastError();
    return nullptr;
}

static int trace_common(sortgpu_accel* acc, const sortgpu_ray* d_rays, uint32_t n, void* d_out, bool any,
                        bool stats, uint64_t* out4, void* stream) {
    if (!acc) {
        set_error(nullptr, "trace: accel is NULL");
        return SORTGPU_ERR_ARG;
    }
    sortgpu_ctx* ctx = acc->ctx;
    if (ctx->host_only) {
        set_error(ctx, "trace: host-only context has no device");
        return SORTGPU_ERR_STATE;
    }
    if (n == 0) {
        if (out4) std::memset(out4, 0, 5 * sizeof(uint64_t));
        return SORTGPU_OK;
    }
    if (!d_rays || !d_out) {
        set_error(ctx, "trace: null ray or output pointer");
        return SORTGPU_ERR_ARG;
    }
    int rc = ctx_bind(ctx);
    if (rc != SORTGPU_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    TraceArgs a{};
    a.tree = acc->view;
    a.rays = d_rays;
    a.n = n;
    a.n_dev = nullptr;
    a.out = d_out;
    const uint32_t slot = acc->slot.fetch_add(1) % kCounterSlots;
    a.counter = acc->d_counters + slot * kCounterStride;
    a.error = acc->d_error;
    a.stats = nullptr;
    apply_tuning(ctx, a);
    if (stats) {
        hipError_t e = hipMemsetAsync(acc->d_stats, 0, 5 * sizeof(unsigned long long), s);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMemsetAsync(stats)");
        a.stats = acc->d_stats;
    }
    // traversal order (results do not depend on it; the counter launch keeps index order)
    RayOrder ro;
    if (!stats) a.order = traversal_order(acc, d_rays, n, s, ro);
    hipError_t e = launch_trace(a, any, stats, s, ctx->num_cus);
    if (ro.order) release_order(ro, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "trace launch");
    if (stats) {
        unsigned long long h[5];
        e = hipMemcpyAsync(h, acc->d_stats, sizeof(h), hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync(stats)");
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
        for (int k = 0; k < 5; k++) out4[k] = h[k];
        return report_flags(ctx, acc->d_error, s, "trace_counters", nullptr);
    }
    return SORTGPU_OK;
}

int sortgpu_trace_closest(sortgpu_accel* acc, const sortgpu_ray* d_rays, uint32_t n, sortgpu_hit* d_hits,
                          void* stream) {
    return trace_common(acc, d_rays, n, d_hits, false, false, nullptr, stream);
}

// sortgpu_trace_closest without any allocation: the traversal order of a large batch is sorted in
// the library's file-scope device array (order.hip) instead of pool memory. The work counter and
// error word are the accelerator's own, allocated when it was built.
int sortgpu_trace_closest_noalloc(sortgpu_accel* acc, const sortgpu_ray* d_rays, uint32_t n, sortgpu_hit* d_hits,
                                  void* stream) {
    if (!acc) {
        set_error(nullptr, "trace_closest_noalloc: accel is NULL");
        return SORTGPU_ERR_ARG;
    }
    sortgpu_ctx* ctx = acc->ctx;
    if (ctx->host_only) {
        set_error(ctx, "trace_closest_noalloc: host-only context has no device");
        return SORTGPU_ERR_STATE;
    }
    if (n > kScratchRays) {
        set_error(ctx, "trace_closest_noalloc: %u rays exceed the no-allocation limit of %u per call; split the batch",
                  n, kScratchRays);
        return SORTGPU_ERR_ARG;
    }
    if (n == 0) return SORTGPU_OK;
    if (!d_rays || !d_hits) {
        set_error(ctx, "trace_closest_noalloc: null ray or output pointer");
        return SORTGPU_ERR_ARG;
    }
    int rc = ctx_bind(ctx);
    if (rc != SORTGPU_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    TraceArgs a{};
    a.tree = acc->view;
    a.rays = d_rays;
    a.n = n;
    a.out = d_hits;
    a.counter = acc->d_counters + (acc->slot.fetch_add(1) % kCounterSlots) * kCounterStride;
    a.error = acc->d_error;
    apply_tuning(ctx, a);
    const uint32_t reorder = ctx->tuning.reorder;
    if (reorder == 2 || (reorder != 1 && n >= kReorderMinRays)) {
        void* base = nullptr;
        hipError_t e = scratch_base(ctx->device, &base);
        if (e != hipSuccess) return hip_fail(ctx, e, "trace_closest_noalloc: scratch address");
        e = order_rays_scratch(base, d_rays, n, acc->view.bbox, s, &a.order);
        if (e != hipSuccess) return hip_fail(ctx, e, "trace_closest_noalloc: traversal order");
    }
    hipError_t e = launch_trace(a, false, false, s, ctx->num_cus);
    if (e != hipSuccess) return hip_fail(ctx, e, "trace_closest_noalloc launch");
    return SORTGPU_OK;
}

int sortgpu_trace_any(sortgpu_accel* acc, const sortgpu_ray* d_rays, uint32_t n, uint8_t* d_occluded,
                      void* stream) {
    return trace_common(acc, d_rays, n, d_occluded, true, false, nullptr, stream);
}

}  // extern "C"

namespace {
// Per-thread staging for the single-ray queries: a stream and 64 bytes of device memory (ray,
// hit, occlusion byte) per device the thread queries, created on its first single-ray call and
// released when the thread ends. SORT's worker threads each get their own, so concurrent calls
// never share a buffer.
struct RayStage {
    int device = -1;
    hipStream_t s = nullptr;
    unsigned char* d = nullptr;
};
struct RayStages {
    RayStage st[8];
    ~RayStages() {
        for (RayStage& r : st) {
            if (r.device < 0) continue;
            (void)hipSetDevice(r.device);
            if (r.d) (void)hipFree(r.d);
            if (r.s) (void)hipStreamDestroy(r.s);
        }
    }
};
thread_local RayStages t_stages;

int single_ray(sortgpu_accel* acc, const sortgpu_ray* ray, bool any, sortgpu_hit* hit, int* occluded) {
    if (!acc || !ray || (!any && !hit) || (any && !occluded)) {
        set_error(acc ? acc->ctx : nullptr, "trace (single ray): null argument");
        return SORTGPU_ERR_ARG;
    }
    sortgpu_ctx* ctx = acc->ctx;
    if (ctx->host_only) {
        set_error(ctx, "trace (single ray): host-only context has no device");
        return SORTGPU_ERR_STATE;
    }
    int rc = ctx_bind(ctx);
    if (rc != SORTGPU_OK) return rc;
    RayStage* stage = nullptr;
    for (RayStage& r : t_stages.st)
        if (r.device == ctx->device) {
            stage = &r;
            break;
        }
    if (!stage) {
        for (RayStage& r : t_stages.st)
            if (r.device < 0) {
                stage = &r;
                break;
            }
        if (!stage) {
            set_error(ctx, "trace (single ray): one thread queries more than 8 devices");
            return SORTGPU_ERR_STATE;
        }
        hipError_t e = hipStreamCreateWithFlags(&stage->s, hipStreamNonBlocking);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamCreate(single ray)");
        void* p = nullptr;
        if ((e = hipMalloc(&p, 64)) != hipSuccess) {
            (void)hipStreamDestroy(stage->s);
            stage->s = nullptr;
            return hip_fail(ctx, e, "hipMalloc(single ray)");
        }
        stage->d = static_cast<unsigned char*>(p);
        stage->device = ctx->device;
    }
    hipError_t e = hipMemcpyAsync(stage->d, ray, sizeof(sortgpu_ray), hipMemcpyHostToDevice, stage->s);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync(single ray)");
    void* d_out = stage->d + 32;
    if ((rc = trace_common(acc, reinterpret_cast<const sortgpu_ray*>(stage->d), 1, d_out, any, false, nullptr,
                           stage->s)) != SORTGPU_OK)
        return rc;
    uint8_t occ = 0;
    e = any ? hipMemcpyAsync(&occ, d_out, 1, hipMemcpyDeviceToHost, stage->s)
            : hipMemcpyAsync(hit, d_out, sizeof(sortgpu_hit), hipMemcpyDeviceToHost, stage->s);
    if (e == hipSuccess) e = hipStreamSynchronize(stage->s);
    if (e != hipSuccess) return hip_fail(ctx, e, "single ray result");
    if (any) *occluded = occ ? 1 : 0;
    return SORTGPU_OK;
}
}  // namespace

extern "C" {

int sortgpu_trace_closest1(sortgpu_accel* acc, const sortgpu_ray* ray, sortgpu_hit* hit) {
    return single_ray(acc, ray, false, hit, nullptr);
}

int sortgpu_trace_any1(sortgpu_accel* acc, const sortgpu_ray* ray, int* occluded) {
    return single_ray(acc, ray, true, nullptr, occluded);
}

int sortgpu_trace_attenuation(sortgpu_accel* acc, const sortgpu_ray* d_rays, uint32_t n, const float* d_transparency,
                              uint32_t n_transparency, float* d_atten, void* stream) {
    if (!acc) {
        set_error(nullptr, "trace_attenuation: accel is NULL");
        return SORTGPU_ERR_ARG;
    }
    sortgpu_ctx* ctx = acc->ctx;
    if (ctx->host_only) {
        set_error(ctx, "trace_attenuation: host-only context has no device");
        return SORTGPU_ERR_STATE;
    }
    if (n == 0) return SORTGPU_OK;
    if (!d_rays || !d_transparency || !d_atten) {
        set_error(ctx, "trace_attenuation: null ray, transparency or output pointer");
        return SORTGPU_ERR_ARG;
    }
    const uint64_t nprims = acc->idx.size() / 3 + acc->quads.size();
    if (n_transparency < nprims) {
        set_error(ctx, "trace_attenuation: %u transparency records for %llu primitives", n_transparency,
                  (unsigned long long)nprims);
        return SORTGPU_ERR_ARG;
    }
    int rc = ctx_bind(ctx);
    if (rc != SORTGPU_OK) return rc;
    TraceArgs a{};
    a.tree = acc->view;
    a.rays = d_rays;
    a.n = n;
    a.out = d_atten;
    const uint32_t slot = acc->slot.fetch_add(1) % kCounterSlots;
    a.counter = acc->d_counters + slot * kCounterStride;
    a.error = acc->d_error;
    a.transp = d_transparency;
    apply_tuning(ctx, a);
    hipStream_t s = static_cast<hipStream_t>(stream);
    RayOrder ro;
    a.order = traversal_order(acc, d_rays, n, s, ro);
    hipError_t e = launch_attenuation(a, s, ctx->num_cus);
    if (ro.order) release_order(ro, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "attenuation launch");
    return SORTGPU_OK;
}

int sortgpu_trace_multi(sortgpu_accel* acc, const sortgpu_ray* d_rays, uint32_t n, const uint32_t* d_ray_material,
                        const uint32_t* d_tri_material, uint32_t n_tri_material, sortgpu_hit* d_hits,
                        uint32_t* d_counts, void* stream) {
    if (!acc) {
        set_error(nullptr, "trace_multi: accel is NULL");
        return SORTGPU_ERR_ARG;
    }
    sortgpu_ctx* ctx = acc->ctx;
    if (ctx->host_only) {
        set_error(ctx, "trace_multi: host-only context has no device");
        return SORTGPU_ERR_STATE;
    }
    if (n == 0) return SORTGPU_OK;
    if (!d_rays || !d_ray_material || !d_hits || !d_counts || (!d_tri_material && !acc->idx.empty())) {
        set_error(ctx, "trace_multi: null ray, material, hit or count pointer");
        return SORTGPU_ERR_ARG;
    }
    const uint64_t ntris = acc->idx.size() / 3;
    if (n_tri_material < ntris) {
        set_error(ctx, "trace_multi: %u triangle materials for %llu triangles", n_tri_material,
                  (unsigned long long)ntris);
        return SORTGPU_ERR_ARG;
    }
    int rc = ctx_bind(ctx);
    if (rc != SORTGPU_OK) return rc;
    MultiArgs a{};
    a.tree = acc->view;
    a.rays = d_rays;
    a.n = n;
    a.ray_material = d_ray_material;
    a.tri_material = d_tri_material;
    a.hits = d_hits;
    a.counts = d_counts;
    a.error = acc->d_error;
    hipStream_t s = static_cast<hipStream_t>(stream);
    RayOrder ro;
    a.order = traversal_order(acc, d_rays, n, s, ro);
    hipError_t e = launch_multi(a, s);
    if (ro.order) release_order(ro, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "multi-hit launch");
    return SORTGPU_OK;
}

int sortgpu_accel_errors(sortgpu_accel* acc, void* stream, uint32_t* flags) {
    if (flags) *flags = 0;
    if (!acc) {
        set_error(nullptr, "sortgpu_accel_errors: accel is NULL");
        return SORTGPU_ERR_ARG;
    }
    if (acc->ctx->host_only) return SORTGPU_OK;
    int rc = ctx_bind(acc->ctx);
    if (rc != SORTGPU_OK) return rc;
    return report_flags(acc->ctx, acc->d_error, static_cast<hipStream_t>(stream), "accelerator", flags);
}

int sortgpu_trace_counters(sortgpu_accel* acc, const sortgpu_ray* d_rays, uint32_t n, int mode,
                           sortgpu_hit* d_hits, uint8_t* d_occluded, uint64_t out[5], void* stream) {
    if (!out || (mode != 0 && mode != 1)) {
        set_error(acc ? acc->ctx : nullptr, "trace_counters: bad mode or null out");
        return SORTGPU_ERR_ARG;
    }
    void* d_out = mode == 0 ? static_cast<void*>(d_hits) : static_cast<void*>(d_occluded);
    return trace_common(acc, d_rays, n, d_out, mode == 1, true, out, stream);
}

}  // extern "C"
