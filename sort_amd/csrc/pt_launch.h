// Host-side launch interface of the path-tracer stages.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "pt_types.h"
#include "trace_launch.h"

namespace sortgpu {

// Optional record of every extension ray (bounce 0 = camera ray) with its closest hit, for
// paths of pixels [pix_begin, pix_end): slot ((pixel - pix_begin) * spp + sample) * depth + bounce.
struct RayDumpDev {
    float4* rays;      // 2 float4 per slot, or null: no record
    float4* hits;      // (t, prim bits, u, v) per slot
    uint8_t* valid;    // 1 where the path traced that bounce
    uint32_t pix_begin, pix_end, spp, depth;
};

struct PtLaunch {
    TraceArgs trace;            // counter / error slots and tuning; the single tree when trees == null
    const SceneDev* scenes;     // device scene table
    const WaveDesc* wave;       // device wave descriptor (chunks + request table)
    const TreeView* trees;      // device tree table for multi-scene waves, or null
    uint32_t n_trees;
    uint32_t tree_width;        // node width of the tree table
    uint32_t tree_max_stack;    // deepest stack over the tree table
    PathState ps;
    Queues q;
    uint32_t max_depth;         // deepest request of the wave
    uint32_t paths_cap;         // grid bound: at least the wave's paths (kernels read the exact count)
    uint32_t pix_cap;           // grid bound: at least the wave's film pixels
    int num_cus;
    RayDumpDev dump;            // rays == null: none
    // single-scene form: the scene's transparency per primitive id (SceneDev::transp), or null;
    // multi-scene form: a device table of those pointers per tree index, or null. Set, shadow rays
    // take the transparent-shadow walk and scale their contributions.
    const float* transp;
    const float* const* transps;
    // each bounce's hits are shaded in spatial order when the wave holds at least this many paths
    // (0 = never)
    uint32_t sort_min;
    // stage timing (or null): while *stage_used + 6 <= stage_cap, a bounce records six events from
    // stage_ev[*stage_used] on (before / after its extension traversal, shading, shadow rays)
    hipEvent_t* stage_ev;
    uint32_t stage_cap;
    uint32_t* stage_used;
};

// Bytes of path workspace per path slot (state + queues + hit order), used for HBM budgeting.
constexpr size_t kPathBytes = 4 * 16 + 2 * (32 + 4 + 4) + 16 + 2 * 32 + 2 * 4 + 2 * 4 + 2 * 4;

hipError_t pt_zero_f4(float4* a, uint32_t n, hipStream_t s);
hipError_t pt_bounces(const PtLaunch& L, uint32_t npaths, hipStream_t s);
// raygen + bounce loop + finalize for the chunks of one wave.
hipError_t pt_wave(const PtLaunch& L, hipStream_t s);
// The two halves of pt_wave: the paths (raygen + bounce loop, touching only the workspace) and
// the film update (adding each chunk's samples, in sample order, to its request's film). Waves
// on different streams may overlap their paths; their film updates must run in wave order. The
// paths half launches the same kernels with the same arguments for every wave of a given
// workspace, descriptor address, depth and capacity, so it can be captured once as a graph.
hipError_t pt_wave_paths(const PtLaunch& L, hipStream_t s);
hipError_t pt_wave_film(const PtLaunch& L, hipStream_t s);
hipError_t pt_resolve_film(const float4* film, uint32_t npix, float* rgb, hipStream_t s);
// Caller rays of request 0 of L.wave's table.
hipError_t pt_li_rays(const PtLaunch& L, const float4* rays, const uint32_t* keys, uint32_t n, float* out,
                      hipStream_t s);

}  // namespace sortgpu
