// Traversal order of a ray batch.
//
// A traversal lane reads its own node and primitive records, and what a launch costs is the
// number of such reads that the compute units serve. Lanes of one wave that walk the same nodes
// at the same time share those reads. Incoherent batches (bounce rays from random surface points)
// arrive in no useful order, so before the launch the batch is sorted (stably) by its origin's
// cell on a 128^3 grid over the tree's box, in Morton order (a 21-bit key: the same three 8-bit
// radix passes as a 64^3 grid, and an order that measured 0.6% faster on incoherent batches).
// Direction cells or octants in the key were measured to cost camera batches 4-5% and gain bounce
// batches under 1%, so the key has none: a batch sharing one origin keeps its order. The kernel takes rays in the sorted
// order; each ray's result is written at its own index and does not depend on the order, so
// results are bit-identical with or without it.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <mutex>

#include "trace_launch.h"

namespace sortgpu {
namespace {

// bits of x (< 2^10) moved to every third position
__device__ __forceinline__ uint32_t spread3(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// cell index in [0, cells) of t in [0, 1]; NaN and out-of-range values clamp to an end cell
__device__ __forceinline__ uint32_t cell(float t, float cells) {
    const float c = fminf(fmaxf(t * cells, 0.0f), cells - 1.0f);  // fmaxf(NaN, 0) = 0
    return (uint32_t)c;
}

__global__ __launch_bounds__(256) void k_ray_keys(const float4* __restrict__ rays, uint32_t n, float lx, float ly,
                                                  float lz, float sx, float sy, float sz, uint32_t* __restrict__ keys,
                                                  uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 q0 = rays[2 * (size_t)i];
    keys[i] = spread3(cell((q0.x - lx) * sx, 128.f)) | (spread3(cell((q0.y - ly) * sy, 128.f)) << 1) |
              (spread3(cell((q0.z - lz) * sz, 128.f)) << 2);
    vals[i] = i;
}

constexpr int kKeyBits = 21;

// Multi-scene form: the scene index above the origin cell in that scene's own box, so the rays of
// one scene stay together and, within it, walk in cell order. The cell takes cell_bits (a multiple
// of 3: kKeyBits, or fewer when the scene index needs more of the 32 bits). Out-of-range scene ids
// sort last (their rays miss whatever the order).
__global__ __launch_bounds__(256) void k_ray_keys_multi(const float4* __restrict__ rays,
                                                        const uint32_t* __restrict__ ray_scene, uint32_t n,
                                                        const TreeView* __restrict__ views, uint32_t n_views,
                                                        uint32_t cell_bits, uint32_t* __restrict__ keys,
                                                        uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t sid = ray_scene[i];
    uint32_t key = 0xffffffffu;
    if (sid < n_views) {
        const float* b = views[sid].bbox;
        const float4 q0 = rays[2 * (size_t)i];
        float ext[3];
        for (int k = 0; k < 3; k++) {
            const float w = b[3 + k] - b[k];
            ext[k] = w > 0.0f ? 1.0f / w : 0.0f;
        }
        const float cells = (float)(1u << (cell_bits / 3));
        key = (sid << cell_bits) | spread3(cell((q0.x - b[0]) * ext[0], cells)) |
              (spread3(cell((q0.y - b[1]) * ext[1], cells)) << 1) | (spread3(cell((q0.z - b[2]) * ext[2], cells)) << 2);
    }
    keys[i] = key;
    vals[i] = i;
}

}  // namespace

void release_order(RayOrder& ro, hipStream_t s) {
    for (void* p : {static_cast<void*>(ro.keys[0]), static_cast<void*>(ro.keys[1]), static_cast<void*>(ro.vals[0]),
                    static_cast<void*>(ro.vals[1]), ro.temp})
        if (p) (void)hipFreeAsync(p, s);
    ro = RayOrder{};
}

hipError_t order_rays(hipMemPool_t pool, const void* rays, uint32_t n, const float bbox[6], hipStream_t s,
                      RayOrder& ro) {
    ro = RayOrder{};
    const size_t b = (size_t)n * sizeof(uint32_t);
    hipError_t e = hipSuccess;
    for (uint32_t** p : {&ro.keys[0], &ro.keys[1], &ro.vals[0], &ro.vals[1]}) {
        if ((e = hipMallocFromPoolAsync(reinterpret_cast<void**>(p), b, pool, s)) != hipSuccess) {
            release_order(ro, s);
            return e;
        }
    }
    float ext[3];
    for (int k = 0; k < 3; k++) {
        const float w = bbox[3 + k] - bbox[k];
        ext[k] = w > 0.0f ? 1.0f / w : 0.0f;
    }
    hipLaunchKernelGGL(k_ray_keys, dim3((n + 255) / 256), dim3(256), 0, s, static_cast<const float4*>(rays), n, bbox[0],
                       bbox[1], bbox[2], ext[0], ext[1], ext[2], ro.keys[0], ro.vals[0]);
    if ((e = hipGetLastError()) != hipSuccess) {
        release_order(ro, s);
        return e;
    }
    hipcub::DoubleBuffer<uint32_t> kb(ro.keys[0], ro.keys[1]), vb(ro.vals[0], ro.vals[1]);
    size_t tb = 0;
    if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kb, vb, (int)n, 0, kKeyBits, s)) != hipSuccess ||
        (e = hipMallocFromPoolAsync(&ro.temp, tb ? tb : 1, pool, s)) != hipSuccess ||
        (e = hipcub::DeviceRadixSort::SortPairs(ro.temp, tb, kb, vb, (int)n, 0, kKeyBits, s)) != hipSuccess) {
        release_order(ro, s);
        return e;
    }
    ro.order = vb.Current();
    return hipSuccess;
}

hipError_t order_rays_multi(hipMemPool_t pool, const void* rays, const uint32_t* ray_scene, uint32_t n,
                            const TreeView* views, uint32_t n_views, hipStream_t s, RayOrder& ro) {
    ro = RayOrder{};
    int sid_bits = 0;
    while ((1ull << sid_bits) <= n_views) sid_bits++;  // room for every id and the out-of-range key
    if (sid_bits > 29) return hipErrorInvalidValue;  // leaves a 2^3 grid at least
    const int cell_bits = (32 - sid_bits) / 3 * 3 < kKeyBits ? (32 - sid_bits) / 3 * 3 : kKeyBits;
    const int bits = cell_bits + sid_bits;
    const size_t b = (size_t)n * sizeof(uint32_t);
    hipError_t e = hipSuccess;
    for (uint32_t** p : {&ro.keys[0], &ro.keys[1], &ro.vals[0], &ro.vals[1]}) {
        if ((e = hipMallocFromPoolAsync(reinterpret_cast<void**>(p), b, pool, s)) != hipSuccess) {
            release_order(ro, s);
            return e;
        }
    }
    hipLaunchKernelGGL(k_ray_keys_multi, dim3((n + 255) / 256), dim3(256), 0, s, static_cast<const float4*>(rays),
                       ray_scene, n, views, n_views, (uint32_t)cell_bits, ro.keys[0], ro.vals[0]);
    if ((e = hipGetLastError()) != hipSuccess) {
        release_order(ro, s);
        return e;
    }
    hipcub::DoubleBuffer<uint32_t> kb(ro.keys[0], ro.keys[1]), vb(ro.vals[0], ro.vals[1]);
    size_t tb = 0;
    // out-of-range ids carry the all-ones key: sorting `bits` bits keeps them last
    if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kb, vb, (int)n, 0, bits < 32 ? bits : 32, s)) !=
            hipSuccess ||
        (e = hipMallocFromPoolAsync(&ro.temp, tb ? tb : 1, pool, s)) != hipSuccess ||
        (e = hipcub::DeviceRadixSort::SortPairs(ro.temp, tb, kb, vb, (int)n, 0, bits < 32 ? bits : 32, s)) !=
            hipSuccess) {
        release_order(ro, s);
        return e;
    }
    ro.order = vb.Current();
    return hipSuccess;
}

// ---- traversal order without allocation -------------------------------------------------------
// A file-scope device array holds the order buffers (keys and indices, double-buffered, and the
// radix sort's workspace) for batches of up to kScratchRays rays, so callers that cannot allocate
// device memory (code under graph capture, serving loops, harnesses that forbid hipMalloc) get the
// same traversal order. Each device holds its own copy for the life of the process; every call on
// a device shares it, so such calls that overlap on one device race.
namespace {
constexpr size_t kScratchArray = (size_t)kScratchRays * sizeof(uint32_t);
constexpr size_t kScratchTemp = kScratchBytes - 4 * kScratchArray;
static_assert(kScratchBytes > 4 * kScratchArray, "scratch must hold the four order arrays and a workspace");
static_assert(kScratchArray % 256 == 0, "order arrays stay 256-byte aligned");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

hipError_t scratch_base(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        // the device must be current: each device holds the array at its own address
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

hipError_t order_rays_scratch(void* base, const void* rays, uint32_t n, const float bbox[6], hipStream_t s,
                              const uint32_t** order) {
    *order = nullptr;
    if (!base || n > kScratchRays) return hipErrorInvalidValue;
    unsigned char* b = static_cast<unsigned char*>(base);
    uint32_t* keys0 = reinterpret_cast<uint32_t*>(b);
    uint32_t* keys1 = reinterpret_cast<uint32_t*>(b + kScratchArray);
    uint32_t* vals0 = reinterpret_cast<uint32_t*>(b + 2 * kScratchArray);
    uint32_t* vals1 = reinterpret_cast<uint32_t*>(b + 3 * kScratchArray);
    void* temp = b + 4 * kScratchArray;
    float ext[3];
    for (int k = 0; k < 3; k++) {
        const float w = bbox[3 + k] - bbox[k];
        ext[k] = w > 0.0f ? 1.0f / w : 0.0f;
    }
    hipcub::DoubleBuffer<uint32_t> kb(keys0, keys1), vb(vals0, vals1);
    size_t tb = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kb, vb, (int)n, 0, kKeyBits, s);
    if (e != hipSuccess) return e;
    if (tb > kScratchTemp) return hipErrorInvalidValue;  // the workspace does not fit: refuse, never overrun
    hipLaunchKernelGGL(k_ray_keys, dim3((n + 255) / 256), dim3(256), 0, s, static_cast<const float4*>(rays), n, bbox[0],
                       bbox[1], bbox[2], ext[0], ext[1], ext[2], keys0, vals0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipcub::DeviceRadixSort::SortPairs(temp, tb, kb, vb, (int)n, 0, kKeyBits, s)) != hipSuccess) return e;
    *order = vb.Current();
    return hipSuccess;
}

}  // namespace sortgpu
