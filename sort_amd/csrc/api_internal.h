// Internal definitions behind the opaque C handles.
#pragma once
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sortgpu.h"
#include "bvh_build.h"
#include "layout.h"

struct sortgpu_ctx {
    int device = -1;
    bool host_only = true;
    int num_cus = 0;
    size_t hbm_bytes = 0;
    char arch[64] = {0};
    sortgpu_tuning tuning{};
    std::mutex mu;
    std::string err;
    std::atomic<int> refs{1};  // the creator's, plus one per object made from the context
    // stream-ordered device memory for traversal orders (created on first use; released with the
    // context once its last allocation is freed)
    hipMemPool_t order_pool = nullptr;
    ~sortgpu_ctx() {
        if (order_pool) (void)hipMemPoolDestroy(order_pool);
    }
};

namespace sortgpu {
// Work-counter ring: each query takes the next slot so queries on different streams that
// share one accelerator never share a counter (unless more than kCounterSlots are in flight).
constexpr uint32_t kCounterSlots = 256;
constexpr uint32_t kCounterStride = 16;  // 64 B apart
}  // namespace sortgpu

struct sortgpu_accel {
    sortgpu_ctx* ctx = nullptr;
    sortgpu::BuiltTree host;
    sortgpu::TreeView view;
    void* d_nodes = nullptr;
    sortgpu::PrimRec* d_prims = nullptr;
    sortgpu::QuadRec* d_quads = nullptr;
    uint32_t* d_counters = nullptr;
    uint32_t* d_error = nullptr;
    unsigned long long* d_stats = nullptr;
    std::atomic<uint32_t> slot{0};
    size_t device_bytes = 0;
    // host copies of the input geometry (shading and hit reconstruction need it by primitive id)
    std::vector<float> pos;
    std::vector<uint32_t> idx;
    std::vector<sortgpu_quad> quads;
};

namespace sortgpu {
void set_error(sortgpu_ctx* ctx, const char* fmt, ...);
int hip_fail(sortgpu_ctx* ctx, hipError_t e, const char* what);
int ctx_bind(sortgpu_ctx* ctx);
// Reference counting of contexts: every object made from a context holds one reference.
void ctx_acquire(sortgpu_ctx* ctx);
void ctx_release(sortgpu_ctx* ctx);
int upload_accel(sortgpu_accel* acc);
// Copies the context's tuning knobs into a launch.
void apply_tuning(const sortgpu_ctx* ctx, struct TraceArgs& a);
void free_accel_device(sortgpu_accel* acc);
// Batches at least this large are traced in coherence order by default (smaller ones do not pay
// for the sort).
constexpr uint32_t kReorderMinRays = 1u << 18;
// Waves of at least this many paths shade each bounce's hits in spatial order (tuning shade_order:
// 0 = this default, 1 = never (0 returned), 2 = always).
inline uint32_t shade_sort_min(const sortgpu_tuning& t) {
    return t.shade_order == 1 ? 0u : t.shade_order == 2 ? 1u : (1u << 18);
}
// The context's stream-ordered pool for traversal-order buffers (made on first use), or null.
hipMemPool_t order_pool(sortgpu_ctx* ctx);
// Takes (and clears) the error flags at d_flags after the work queued on `s`; returns SORTGPU_OK
// when none was raised, else SORTGPU_ERR_STATE with a message naming each flag ("who: ...").
int report_flags(sortgpu_ctx* ctx, uint32_t* d_flags, hipStream_t s, const char* who, uint32_t* flags_out);
}  // namespace sortgpu
