// Batched form: the records of many accelerators concatenated into one device allocation, with a
// device table of their tree views, traced by one persistent launch (k_trace's MULTI form).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/sortgpu.h"
#include "api_internal.h"
#include "trace_launch.h"

using namespace sortgpu;

struct sortgpu_pool {
    sortgpu_ctx* ctx = nullptr;
    uint32_t n = 0, width = 8, max_stack = 1;
    void* d_block = nullptr;       // all node / primitive / quad records
    TreeView* d_views = nullptr;   // n tree views
    uint32_t* d_counters = nullptr;
    uint32_t* d_error = nullptr;
    size_t bytes = 0;
    std::atomic<uint32_t> slot{0};
};

namespace {
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace

extern "C" {

int sortgpu_pool_create(sortgpu_ctx* ctx, sortgpu_accel* const* accels, uint32_t n, sortgpu_pool** out) {
    if (!ctx || !out || (n && !accels)) {
        set_error(ctx, "sortgpu_pool_create: null argument");
        return SORTGPU_ERR_ARG;
    }
    *out = nullptr;
    if (ctx->host_only) {
        set_error(ctx, "sortgpu_pool_create: host-only context");
        return SORTGPU_ERR_STATE;
    }
    uint32_t width = 0, max_stack = 1;
    size_t total = 0;
    for (uint32_t i = 0; i < n; i++) {
        const sortgpu_accel* a = accels[i];
        if (!a) {
            set_error(ctx, "sortgpu_pool_create: accelerator %u is NULL", i);
            return SORTGPU_ERR_ARG;
        }
        if (a->host.valid) {
            if (width && a->host.width != width) {
                set_error(ctx, "sortgpu_pool_create: mixed node widths (%u and %u)", width, a->host.width);
                return SORTGPU_ERR_ARG;
            }
            width = a->host.width;
        }
        max_stack = std::max(max_stack, a->host.max_stack);
        total += align256(a->host.node_bytes.size()) + align256(a->host.prims.size() * sizeof(PrimRec)) +
                 align256(a->host.quads.size() * sizeof(QuadRec));
    }
    if (!width) width = 8;
    if (max_stack > trace_max_stack_supported()) {
        set_error(ctx, "sortgpu_pool_create: a tree needs %u stack entries (limit %u)", max_stack,
                  trace_max_stack_supported());
        return SORTGPU_ERR_ARG;
    }
    int rc = ctx_bind(ctx);
    if (rc) return rc;
    sortgpu_pool* p = new (std::nothrow) sortgpu_pool();
    if (!p) return SORTGPU_ERR_OOM;
    p->ctx = ctx;
    ctx_acquire(ctx);
    p->n = n;
    p->width = width;
    p->max_stack = max_stack;
    hipError_t e;
    auto fail = [&](hipError_t err, const char* what) {
        int r = hip_fail(ctx, err, what);
        sortgpu_pool_destroy(p);
        return r;
    };
    if ((e = hipMalloc(&p->d_block, std::max<size_t>(total, 256))) != hipSuccess) return fail(e, "hipMalloc(pool)");
    if ((e = hipMalloc(reinterpret_cast<void**>(&p->d_views), std::max<size_t>(n, 1) * sizeof(TreeView))) != hipSuccess)
        return fail(e, "hipMalloc(pool views)");
    if ((e = hipMalloc(reinterpret_cast<void**>(&p->d_counters), kCounterSlots * kCounterStride * sizeof(uint32_t))) !=
        hipSuccess)
        return fail(e, "hipMalloc(pool counters)");
    if ((e = hipMalloc(reinterpret_cast<void**>(&p->d_error), 16)) != hipSuccess) return fail(e, "hipMalloc(error)");
    if ((e = hipMemset(p->d_counters, 0, kCounterSlots * kCounterStride * sizeof(uint32_t))) != hipSuccess ||
        (e = hipMemset(p->d_error, 0, 16)) != hipSuccess)
        return fail(e, "hipMemset(pool)");
    std::vector<TreeView> views(n);
    char* base = static_cast<char*>(p->d_block);
    size_t off = 0;
    for (uint32_t i = 0; i < n; i++) {
        const BuiltTree& h = accels[i]->host;
        TreeView v;
        std::memset(&v, 0, sizeof(v));
        const size_t nb = h.node_bytes.size(), pb = h.prims.size() * sizeof(PrimRec), qb = h.quads.size() * sizeof(QuadRec);
        v.nodes = base + off;
        if (nb && (e = hipMemcpy(base + off, h.node_bytes.data(), nb, hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "hipMemcpy(pool nodes)");
        off += align256(nb);
        v.prims = reinterpret_cast<const PrimRec*>(base + off);
        if (pb && (e = hipMemcpy(base + off, h.prims.data(), pb, hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "hipMemcpy(pool prims)");
        off += align256(pb);
        v.quads = reinterpret_cast<const QuadRec*>(base + off);
        if (qb && (e = hipMemcpy(base + off, h.quads.data(), qb, hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "hipMemcpy(pool quads)");
        off += align256(qb);
        v.root = h.root_ref;
        v.valid = h.valid ? 1u : 0u;
        v.max_stack = h.max_stack;
        v.width = h.width;
        std::memcpy(v.bbox, h.bbox, sizeof(v.bbox));
        views[i] = v;
    }
    if (n && (e = hipMemcpy(p->d_views, views.data(), n * sizeof(TreeView), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "hipMemcpy(pool views)");
    p->bytes = total;
    *out = p;
    return SORTGPU_OK;
}

void sortgpu_pool_destroy(sortgpu_pool* p) {
    if (!p) return;
    ctx_bind(p->ctx);
    void* ptrs[] = {p->d_block, p->d_views, p->d_counters, p->d_error};
    for (void* x : ptrs)
        if (x) hipFree(x);
    sortgpu_ctx* ctx = p->ctx;
    delete p;
    ctx_release(ctx);
}

int sortgpu_pool_info(const sortgpu_pool* p, uint64_t out[4]) {
    if (!p || !out) return SORTGPU_ERR_ARG;
    out[0] = p->n;
    out[1] = p->width;
    out[2] = p->bytes;
    out[3] = p->max_stack;
    return SORTGPU_OK;
}

int sortgpu_pool_errors(sortgpu_pool* p, void* stream, uint32_t* flags) {
    if (flags) *flags = 0;
    if (!p) {
        set_error(nullptr, "sortgpu_pool_errors: pool is NULL");
        return SORTGPU_ERR_ARG;
    }
    int rc = ctx_bind(p->ctx);
    if (rc) return rc;
    return report_flags(p->ctx, p->d_error, static_cast<hipStream_t>(stream), "pool", flags);
}

int sortgpu_pool_trace(sortgpu_pool* p, const sortgpu_ray* d_rays, const uint32_t* d_ray_scene, uint32_t n, int any,
                       void* d_out, void* stream) {
    if (!p) {
        set_error(nullptr, "sortgpu_pool_trace: pool is NULL");
        return SORTGPU_ERR_ARG;
    }
    if (n == 0) return SORTGPU_OK;
    if (!d_rays || !d_ray_scene || !d_out) {
        set_error(p->ctx, "sortgpu_pool_trace: null pointer");
        return SORTGPU_ERR_ARG;
    }
    int rc = ctx_bind(p->ctx);
    if (rc) return rc;
    TraceArgs a{};
    a.rays = d_rays;
    a.n = n;
    a.out = d_out;
    a.counter = p->d_counters + (p->slot.fetch_add(1) % kCounterSlots) * kCounterStride;
    a.error = p->d_error;
    a.trees = p->d_views;
    a.ray_scene = d_ray_scene;
    a.n_trees = p->n;
    a.width = p->width;
    a.max_stack = p->max_stack;
    apply_tuning(p->ctx, a);
    hipStream_t s = static_cast<hipStream_t>(stream);
    // traversal order of a large batch: by scene, then by origin cell in that scene's box (results
    // do not depend on it; an order that cannot be had only costs speed)
    RayOrder ro;
    const uint32_t reorder = p->ctx->tuning.reorder;
    if (reorder == 2 || (reorder != 1 && n >= kReorderMinRays && n <= 0x7fffffffu)) {
        hipMemPool_t mp = order_pool(p->ctx);
        if (mp && order_rays_multi(mp, d_rays, d_ray_scene, n, p->d_views, p->n, s, ro) == hipSuccess)
            a.order = ro.order;
        else
            (void)hipGetLastError();
    }
    hipError_t e = launch_trace(a, any != 0, false, s, p->ctx->num_cus);
    if (ro.order) release_order(ro, s);
    if (e != hipSuccess) return hip_fail(p->ctx, e, "pool trace launch");
    return SORTGPU_OK;
}

}  // extern "C"
