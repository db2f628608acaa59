// Path-tracer data shared by the host drivers (single render, batched Li, served waves) and the
// kernels.
#pragma once
#include <stdint.h>

#include "layout.h"
#include "shading.h"

namespace sortgpu {

// One renderable scene resident in HBM: the tree plus everything shading reads.
struct SceneDev {
    TreeView tree;
    const uint32_t* idx;       // ntris x 3
    const float* pos;          // nverts x 3
    const float* nrm;          // nverts x 3 shading normals, or null
    const float* tan;          // nverts x 3 tangents, or null
    const uint32_t* tri_mat;   // ntris, or null (material 0)
    const MatRec* mats;        // nmats + 1 (the last is the default white Lambert)
    const LightRec* lights;    // nlights
    const float* light_cdf;    // nlights + 1
    const int32_t* quad_light; // nquads: light index of each quad
    uint32_t ntris, nquads, nlights, nmats;
    int32_t env_light;         // index of the environment light, or -1
    V3 env_L;
    float clamp;               // > 0: per-sample radiance clamp
    // Material::EvaluateTransparency per primitive id (r, g, b, has-transparency), or null when no
    // material is transparent: then shadow rays are any-hit queries, else transparent-shadow walks
    const float* transp;
};

struct CamDev {
    V3 eye, right, up, fwd;
    float x_scale, y_scale;  // 1/tan(fov/2) / aspect, 1/tan(fov/2)
    uint32_t width, height;
};

// One render request: which scene, how to shoot primary rays, where its samples accumulate.
struct ReqDev {
    uint32_t scene;      // index into the scene (and tree) tables
    uint32_t seed;
    uint32_t max_depth;
    uint32_t npix;
    CamDev cam;
    float4* film;        // npix accumulators: summed rgb of valid samples, valid count
    float clamp;         // > 0: each sample's radiance is clamped to [0, clamp] (its scene's)
};

// A wave's work: chunks of whole samples of requests. Paths [path_base, path_base + npix*nsamp)
// belong to chunk k (pixel-major within a sample); pixels [pix_base, pix_base + npix) of the
// finalize grid. Passed by value as a kernel argument.
constexpr int kMaxChunks = 32;
struct ChunkDev {
    uint32_t req, sample0, nsamp, npix, path_base, pix_base;
};
struct ChunkTable {
    uint32_t n, total_paths, total_pix, pad;
    ChunkDev c[kMaxChunks];
};

// Everything a wave's kernels read about its work, in device memory: the chunks and the request
// table they index. A wave's launches take only its address, so a captured wave graph replays for
// any wave whose descriptor has been copied to the same address.
struct WaveDesc {
    ChunkTable ct;
    ReqDev reqs[kMaxChunks];
};

// Path slots of one wave (fixed for the wave; index = path id): one 64-byte record per path, so
// the shading stage reads and writes a path's whole state in one 128-byte line (two paths to a
// line) instead of gathering seven arrays. Record of path p, float4 index 4p + k:
//   k = 0  throughput (xyz); w: pixel index (the RNG key), as bits
//   k = 1  radiance so far (xyz); w: sample index (the RNG counter), as bits
//   k = 2  pending light-sampled NEE contribution of the last vertex (xyz); w: request index, as bits
//   k = 3  pending BSDF-sampled NEE contribution of the last vertex (xyz); w: 0
// Shadow-ray slot s = 2p + j is the contribution at 4p + 2 + j; clearing or scaling one keeps its w.
struct PathState {
    float4* rec;
};
__host__ __device__ inline size_t path_slot(uint32_t s) { return 4 * (size_t)(s >> 1) + 2 + (s & 1u); }

struct Queues {
    float4* ray[2];     // 2 x float4 per entry
    uint32_t* pid[2];
    uint32_t* rsc[2];   // scene of each extension entry (multi-scene traversal)
    float4* hit;        // per entry of the current queue
    float4* sray;       // shadow rays
    uint32_t* ssid;     // shadow slot: 2 * path + k
    uint32_t* ssc;      // scene of each shadow entry
    uint32_t* count;    // [0],[1] extension queue counts, [2] shadow count
    unsigned long long* totals;  // [0] extension rays, [1] shadow rays
    // spatial order of a bounce's hits (bucket sort): bucket per entry, the order, per-block bucket
    // counts (kSortBlocks x kSortBuckets) and bucket bases (kSortBuckets)
    uint32_t* skey;
    uint32_t* sperm;
    uint32_t* shist;
    uint32_t* sbase;
};

// Bucket sort of a bounce's hits: 4096 buckets (a 16^3 grid over the scene's box; in multi-scene
// waves 8 scene classes x 8^3 cells), counted by kSortBlocks workgroups over contiguous slices.
constexpr uint32_t kSortBuckets = 4096;
constexpr uint32_t kSortBlocks = 512;

}  // namespace sortgpu
