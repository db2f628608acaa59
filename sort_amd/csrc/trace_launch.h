// Host-visible launch interface of the traversal kernels.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "layout.h"

namespace sortgpu {

struct TraceArgs {
    TreeView tree;                 // single-scene form
    const void* rays;              // n x 32 B ray records (device)
    uint32_t n;                    // ray count, or the capacity when n_dev is set
    const uint32_t* n_dev;         // optional device-side ray count (read at kernel start)
    void* out;                     // closest: n x 16 B hits; any: n x u8 (device)
    uint32_t* counter;             // work counter (device), reset by the launch
    uint32_t* error;               // error flags (device): kFlag* bits
    unsigned long long* stats;     // 5 counters (device), only with stats
    // multi-scene form: trees != null selects it
    const TreeView* trees;         // device table of tree views
    const uint32_t* ray_scene;     // per ray: index into trees (out-of-range ids miss, error bit1)
    uint32_t n_trees;
    uint32_t width;                // node width shared by every tree of the table
    uint32_t max_stack;            // deepest worst-case stack over the table
    // tuning (0 = default)
    uint32_t refill;               // idle lanes that trigger a refill (default 16)
    uint32_t lds_entries;          // LDS stack entries per lane: 8 or 16 (default 16)
    uint32_t blocks_per_cu;        // cap on resident workgroups per CU (default: occupancy limit)
    uint32_t leaf_batch;           // lanes holding a leaf that end a node phase (default 32)
    uint32_t leaf_min;             // a leaf phase ends below this many lanes holding a leaf (default 1)
    // optional traversal order: the k-th ray taken from the work counter is rays[order[k]] (a
    // permutation of 0..n-1, device), its result goes to out[order[k]]; null = index order.
    const uint32_t* order = nullptr;
    // transparent-shadow attenuation only: per primitive id, 4 floats (transparency r, g, b, and
    // 1 when the primitive's material has transparency, else 0); in the multi-scene form, a device
    // table of such pointers per tree index instead (a null entry: every primitive opaque)
    const float* transp = nullptr;
    const float* const* transps = nullptr;
    // transparent-shadow attenuation only: ray i's result goes to float4 index i * out_stride of out
    // (0 or 1: packed; 2: over the first half of ray i's own 32-byte record when out == rays)
    uint32_t out_stride = 0;
    // any-hit only, for the path tracer's shadow rays: with occ_slot set, an occluded ray i clears
    // the colour (x, y, z) of its path's contribution slot, float4 occ_rec[4 (s / 2) + 2 + s % 2]
    // for s = occ_slot[i], keeping its w, and nothing is written to out; an unoccluded ray writes
    // nothing
    const uint32_t* occ_slot = nullptr;
    float4* occ_rec = nullptr;
};

// Launches the closest (any = false) or any-hit traversal on `stream`.
hipError_t launch_trace(const TraceArgs& a, bool any, bool stats, hipStream_t stream, int num_cus);
// Launches the transparent-shadow attenuation walk (single tree with a.transp, or the multi-scene
// form with a.trees and a.transps): out receives one float4 per ray (attenuation r, g, b, and the
// number of surfaces passed as a float), at stride a.out_stride.
hipError_t launch_attenuation(const TraceArgs& a, hipStream_t stream, int num_cus);
// Most surfaces one attenuation walk passes before it gives up (error bit 2).
constexpr uint32_t kMaxAttenuationSteps = 4096;

// BSSRDF multi-hit query (multihit.hip): per ray four (t, prim, u, v) records and a count.
struct MultiArgs {
    TreeView tree;
    const void* rays;               // n x 32 B ray records (device)
    uint32_t n;
    const uint32_t* ray_material;   // per ray: the material its hits must have
    const uint32_t* tri_material;   // per triangle id: its material
    void* hits;                     // n x 4 x 16 B records
    uint32_t* counts;               // n records filled
    uint32_t* error;                // bit0: stack overflow
    const uint32_t* order = nullptr; // optional traversal order (as TraceArgs::order)
};
hipError_t launch_multi(const MultiArgs& a, hipStream_t stream);

// Traversal order of a ray batch (order.hip): ray indices sorted by the origin's cell in the tree
// box, so that a wave's lanes walk nearby nodes. Buffers come from `pool`, stream-ordered on `s`;
// release_order frees them on `s` after the launch that reads `order`.
struct RayOrder {
    uint32_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    void* temp = nullptr;
    const uint32_t* order = nullptr;
};
hipError_t order_rays(hipMemPool_t pool, const void* rays, uint32_t n, const float bbox[6], hipStream_t s,
                      RayOrder& ro);
void release_order(RayOrder& ro, hipStream_t s);
// The same for rays of many trees: sorted by (tree index, origin cell in that tree's box).
hipError_t order_rays_multi(hipMemPool_t pool, const void* rays, const uint32_t* ray_scene, uint32_t n,
                            const TreeView* views, uint32_t n_views, hipStream_t s, RayOrder& ro);
// Traversal order without allocation (the no-allocation entry point): buffers for batches of up
// to kScratchRays rays live in a file-scope device array of kScratchBytes per device.
constexpr uint32_t kScratchRays = 1u << 22;
constexpr size_t kScratchBytes = (size_t)80 << 20;  // 4 x 16 MiB order arrays + 16 MiB sort workspace
constexpr int kMaxScratchDevices = 64;
// Device address of the calling thread's current device's array (taken once per device with
// hipGetSymbolAddress; `device` must be current).
hipError_t scratch_base(int device, void** out);
// order_rays on the array at `base` (n <= kScratchRays); *order points into it.
hipError_t order_rays_scratch(void* base, const void* rays, uint32_t n, const float bbox[6], hipStream_t s,
                              const uint32_t** order);
// Error flags the traversal kernels raise (atomicOr into a 16-byte device word block).
constexpr uint32_t kFlagStackOverflow = 1u;  // a push found the stack full: that ray's result is incomplete
constexpr uint32_t kFlagBadScene = 2u;       // multi-scene form: a ray named a tree outside the table
constexpr uint32_t kFlagAttenCap = 4u;       // an attenuation walk passed kMaxAttenuationSteps surfaces
// Takes the flags word d_flags[0] (clearing it) after the work queued on `stream`; d_flags[1]
// is the staging word. Synchronises `stream`.
hipError_t take_flags(uint32_t* d_flags, hipStream_t stream, uint32_t* out);
// Deepest stack (entries) the kernels can hold.
uint32_t trace_max_stack_supported();

}  // namespace sortgpu
