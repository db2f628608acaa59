// Host-side helpers shared by the single-scene render API and the serving engine: device
// allocation lists, the path workspace, the request-table ring, scene images (everything a
// renderable scene puts in HBM, prepared on the host) and the camera basis.
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

// Wave descriptors of waves in flight: a ring of slots, each a pinned host staging descriptor and
// a device copy. push() waits until the slot's previous wave has finished with it, stages `w`
// (its chunk table and the first w.ct.n requests), copies it on `s` to `dst` (or, when null, to
// the slot's own device copy) and returns the device address; done() records that the wave that
// reads the slot has been enqueued on `s`.
struct ReqRing {
    static constexpr int kSlots = 8;
    WaveDesc* host = nullptr;
    WaveDesc* dev = nullptr;
    hipEvent_t ev[kSlots] = {};
    uint32_t next = 0, last = 0;
    int init(sortgpu_ctx* ctx);
    void release();
    int push(sortgpu_ctx* ctx, const WaveDesc& w, hipStream_t s, WaveDesc* dst, const WaveDesc** d_desc);
    int done(sortgpu_ctx* ctx, hipStream_t s);
};

// Everything one renderable scene puts in HBM, prepared on the host: optionally its tree records,
// and the shading data (geometry by primitive id, materials + the default white Lambert, light
// records in Scene::genLightDistribution order, the light-pick CDF, the quad -> light map).
struct SceneImage {
    bool with_tree = false;
    TreeView tree{};                    // pointers unset
    std::vector<unsigned char> nodes;
    std::vector<PrimRec> prims;
    std::vector<QuadRec> quads;
    std::vector<uint32_t> idx, tri_mat;
    std::vector<float> pos, nrm, tan, cdf;
    std::vector<float> transp;          // 4 per primitive id when some material is transparent, else empty
    std::vector<MatRec> mats;
    std::vector<LightRec> lights;
    std::vector<int32_t> quad_light;
    SceneDev dev{};                     // counts and environment; pointers unset
    // Bytes of the device block (every array 256-byte aligned).
    size_t block_bytes() const;
    // Copies the arrays into `block` (a device allocation of block_bytes()) and fills `out` with
    // their device addresses; without a tree of its own, out.tree's record pointers stay null
    // for the caller to set.
    // The copies are enqueued on `s`; the image must stay alive until they complete.
    int place(sortgpu_ctx* ctx, void* block, SceneDev& out, hipStream_t s) const;
};

// Validates `desc` against the accelerator's geometry and prepares the scene image (with a copy of
// the accelerator's tree records when with_tree is set). Host-only: no device calls.
int prepare_scene(sortgpu_ctx* ctx, const sortgpu_accel* acc, const sortgpu_scene_desc* desc, bool with_tree,
                  SceneImage& out);

// Camera basis of PerspectiveCamera (camera/perspective.cpp:26-64; math/transform.h).
int make_camera(sortgpu_ctx* ctx, const sortgpu_camera* cam, uint32_t width, uint32_t height, CamDev& out);

inline uint32_t fold_seed(uint64_t seed) { return (uint32_t)(seed ^ (seed >> 32)); }

}  // namespace sortgpu
