// Host-side launch interface of the path-tracer stages.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "pt_types.h"
#include "trace_launch.h"

namespace sortgpu {

struct PtLaunch {
    TraceArgs trace;            // tree, counter and error slots; rays / n / out set per stage
    const SceneDev* scenes;     // device pointer (one scene)
    PathState ps;
    Queues q;
    uint32_t seed;
    uint32_t max_depth;
    int num_cus;
};

hipError_t pt_zero_f4(float4* a, uint32_t n, hipStream_t s);
hipError_t pt_bounces(const PtLaunch& L, uint32_t npaths, hipStream_t s);
hipError_t pt_render_batch(const PtLaunch& L, const CamDev& cam, uint32_t sample0, uint32_t nsamp, float4* film,
                           hipStream_t s);
hipError_t pt_resolve_film(const float4* film, uint32_t npix, float* rgb, hipStream_t s);
hipError_t pt_li_rays(const PtLaunch& L, const float4* rays, const uint32_t* keys, uint32_t n, float* out,
                      hipStream_t s);

}  // namespace sortgpu
