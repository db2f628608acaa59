// Host-side helpers shared by the single-scene render API and the serving engine: device
// allocation lists, the path workspace, scene (shading data) upload and the camera basis.
#pragma once
#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "../../include/sortgpu.h"
#include "api_internal.h"
#include "pt_launch.h"

namespace sortgpu {

// Device allocations owned together; bytes counts what was requested.
struct DeviceAllocs {
    std::vector<void*> ptrs;
    size_t bytes = 0;
    int alloc(sortgpu_ctx* ctx, void** p, size_t n, const char* what);
    int upload(sortgpu_ctx* ctx, const void* src, size_t n, void** dst, const char* what);
    void free_all();
};

// Path slots + queues for up to `cap` paths of one wave.
struct PathWorkspace {
    uint32_t cap = 0;
    PathState ps{};
    Queues q{};
    DeviceAllocs mem;
    int ensure(sortgpu_ctx* ctx, uint32_t paths);
    void release();
    static size_t bytes_for(uint32_t paths);
};

// Request tables of waves in flight: a ring of slots, each a pinned host staging table and its
// device copy. push() waits until the slot's previous wave has finished with it, stages `n`
// records, copies them to the device on `s` and returns the device table; done() records that
// the wave that reads the slot has been enqueued on `s`.
struct ReqRing {
    static constexpr int kSlots = 8;
    ReqDev* host = nullptr;
    ReqDev* dev = nullptr;
    hipEvent_t ev[kSlots] = {};
    uint32_t next = 0, last = 0;
    int init(sortgpu_ctx* ctx);
    void release();
    int push(sortgpu_ctx* ctx, const ReqDev* r, uint32_t n, hipStream_t s, const ReqDev** d_table);
    int done(sortgpu_ctx* ctx, hipStream_t s);
};

// Validates `desc` against the accelerator's geometry and uploads the shading data; `out.tree`
// is the accelerator's own tree view (its device records). Returns a status.
int upload_scene(sortgpu_ctx* ctx, const sortgpu_accel* acc, const sortgpu_scene_desc* desc, SceneDev& out,
                 DeviceAllocs& mem);

// Camera basis of PerspectiveCamera (camera/perspective.cpp:26-64; math/transform.h).
int make_camera(sortgpu_ctx* ctx, const sortgpu_camera* cam, uint32_t width, uint32_t height, CamDev& out);

inline uint32_t fold_seed(uint64_t seed) { return (uint32_t)(seed ^ (seed >> 32)); }

}  // namespace sortgpu
