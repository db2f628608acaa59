// Device residency of a built accelerator: one allocation per record array, uploaded once.
#include <hip/hip_runtime_api.h>

#include <cstring>

#include "api_internal.h"

namespace sortgpu {

void free_accel_device(sortgpu_accel* acc) {
    if (!acc || !acc->ctx || acc->ctx->host_only) return;
    ctx_bind(acc->ctx);
    void* ptrs[] = {acc->d_nodes, acc->d_prims, acc->d_quads, acc->d_counters, acc->d_error, acc->d_stats};
    for (void* p : ptrs)
        if (p) hipFree(p);
    acc->d_nodes = nullptr;
    acc->d_prims = nullptr;
    acc->d_quads = nullptr;
    acc->d_counters = nullptr;
    acc->d_error = nullptr;
    acc->d_stats = nullptr;
}

static int dev_alloc_copy(sortgpu_ctx* ctx, void** dst, const void* src, size_t bytes, const char* what) {
    *dst = nullptr;
    if (bytes == 0) bytes = 16;  // keep every pointer valid and aligned
    hipError_t e = hipMalloc(dst, bytes);
    if (e != hipSuccess) return hip_fail(ctx, e, what);
    if (src) {
        e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    } else {
        e = hipMemset(*dst, 0, bytes);
    }
    if (e != hipSuccess) return hip_fail(ctx, e, what);
    return 0;
}

int upload_accel(sortgpu_accel* acc) {
    sortgpu_ctx* ctx = acc->ctx;
    int rc = ctx_bind(ctx);
    if (rc) return rc;
    const BuiltTree& h = acc->host;
    const size_t nb = h.node_bytes.size(), pb = h.prims.size() * sizeof(PrimRec),
                 qb = h.quads.size() * sizeof(QuadRec);
    if ((rc = dev_alloc_copy(ctx, &acc->d_nodes, nb ? h.node_bytes.data() : nullptr, nb, "hipMalloc(nodes)")) ||
        (rc = dev_alloc_copy(ctx, reinterpret_cast<void**>(&acc->d_prims), h.prims.data(), pb, "hipMalloc(prims)")) ||
        (rc = dev_alloc_copy(ctx, reinterpret_cast<void**>(&acc->d_quads), qb ? h.quads.data() : nullptr, qb,
                             "hipMalloc(quads)")) ||
        (rc = dev_alloc_copy(ctx, reinterpret_cast<void**>(&acc->d_counters), nullptr,
                             kCounterSlots * kCounterStride * sizeof(uint32_t), "hipMalloc(counters)")) ||
        (rc = dev_alloc_copy(ctx, reinterpret_cast<void**>(&acc->d_error), nullptr, 16, "hipMalloc(error)")) ||
        (rc = dev_alloc_copy(ctx, reinterpret_cast<void**>(&acc->d_stats), nullptr, 64, "hipMalloc(stats)"))) {
        free_accel_device(acc);
        return rc;
    }
    acc->device_bytes = nb + pb + qb;
    acc->view.nodes = acc->d_nodes;
    acc->view.prims = acc->d_prims;
    acc->view.quads = acc->d_quads;
    return 0;
}

}  // namespace sortgpu
