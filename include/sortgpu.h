/*
 * sortgpu.h — C ABI of the MI355X-native ray–scene intersection and path-tracing library.
 *
 * This is the drop-in boundary for SORT's spatial accelerator (the `Accelerator` plugin
 * surface: Build / GetIntersect / IsOccluded) and its path-tracing integrator entry (`Li`).
 * Every entry point takes plain pointers and sizes; no C++ or framework types cross it, and
 * no exception crosses it: each call returns SORTGPU_OK (0) or a negative status, with a
 * message available from sortgpu_last_error().
 *
 * Citations name the reference interface each entry replaces (paths relative to SORT's src/).
 *
 * Pointer conventions: arguments named d_* are device pointers (HBM) that the caller owns;
 * arguments without the prefix are host pointers. `stream` is a hipStream_t (NULL = default
 * stream); calls are stream-ordered and return without synchronising unless stated.
 */
#ifndef SORTGPU_H
#define SORTGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (replaces SORT's bool returns + sAssert, core/sassert.h:27-45) ---- */
#define SORTGPU_OK 0
#define SORTGPU_ERR_ARG (-1)    /* invalid argument (null pointer, bad size, bad option)   */
#define SORTGPU_ERR_HIP (-2)    /* a HIP runtime call failed                                */
#define SORTGPU_ERR_OOM (-3)    /* device or host allocation failed                         */
#define SORTGPU_ERR_BUDGET (-4) /* request refused by the HBM admission budget              */
#define SORTGPU_ERR_STATE (-5)  /* object in the wrong state (e.g. query before build)      */

/* ---- records ---- */

/* One ray: origin, tmin, direction, tmax (32 B). Mirrors the algorithmic part of SORT's Ray
 * (math/ray.h:37-93: m_Ori, m_Dir, m_fMin, m_fMax). Directions need not be normalised. */
typedef struct sortgpu_ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} sortgpu_ray;

/* Closest hit (16 B). prim = primitive id in build order (triangles 0..ntris-1, then quads
 * ntris..ntris+nquads-1), 0xFFFFFFFF on a miss (t = FLT_MAX). u, v = barycentric weights of
 * vertices 1 and 2 (simd/simd_triangle.h:296-297); 0 for quads. */
typedef struct sortgpu_hit {
    float t;
    uint32_t prim;
    float u, v;
} sortgpu_hit;

/* Analytic area-light quad (shape/quad.cpp): local plane y = 0, |x| <= size_x/2,
 * |z| <= size_y/2, normal +y; mat = light-to-world 4x4 row-major, inv = its inverse. */
typedef struct sortgpu_quad {
    float mat[16];
    float inv[16];
    float size_x, size_y;
} sortgpu_quad;

/* Build options (Fbvh::Serialize, accel/fast_bvh.h:159-162). width 8 = OBVH, 4 = QBVH.
 * Zero fields take the defaults: width 8, max_depth 28, max_leaf 16 (the Blender exporter's
 * defaults); build_threads 0 = all host cores. */
typedef struct sortgpu_build_opts {
    uint32_t width;
    uint32_t max_depth;
    uint32_t max_leaf;
    uint32_t build_threads;
} sortgpu_build_opts;

typedef struct sortgpu_ctx sortgpu_ctx;
typedef struct sortgpu_accel sortgpu_accel;

/* ---- context ---- */

/* Binds the library to a HIP device. device < 0 creates a host-only context: builds run and
 * can be inspected, but no device memory is used and queries return SORTGPU_ERR_STATE. */
int sortgpu_ctx_create(int device, sortgpu_ctx** out);
void sortgpu_ctx_destroy(sortgpu_ctx* ctx);
/* Last error message of the context (or of the library when ctx is NULL). */
const char* sortgpu_last_error(const sortgpu_ctx* ctx);

/* ---- accelerator: Build (Accelerator::Build, accel/accelerator.h:128; Fbvh::Build,
 *      accel/fast_bvh.hpp:109-136) ----
 * pos: nverts x 3 floats; idx: ntris x 3 vertex indices; quads: nquads records. The SAH
 * 8-wide (or 4-wide) tree is built on the host, flattened into the device layout and
 * uploaded. Input arrays are copied; the caller keeps ownership. An empty scene yields a
 * valid handle whose queries all miss (m_isValid = false, fast_bvh.hpp:112-114). At most
 * 2^30 - 1 primitives (triangles + quads) per tree (SORTGPU_ERR_ARG beyond). */
int sortgpu_accel_build(sortgpu_ctx* ctx, const float* pos, uint32_t nverts, const uint32_t* idx,
                        uint32_t ntris, const sortgpu_quad* quads, uint32_t nquads,
                        const sortgpu_build_opts* opts, sortgpu_accel** out);
void sortgpu_accel_destroy(sortgpu_accel* accel);
/* 1 when the tree holds at least one primitive (Accelerator::m_isValid). */
int sortgpu_accel_valid(const sortgpu_accel* accel);
/* Enlarged scene box used by the root test (core/scene.cpp:145-154): min xyz, max xyz. */
int sortgpu_accel_bbox(const sortgpu_accel* accel, float out6[6]);
/* out[0..7] = width, depth, inner nodes, leaves, primitives, max stack entries,
 *             device bytes (nodes + primitive records), build milliseconds. */
int sortgpu_accel_info(const sortgpu_accel* accel, uint64_t out[8]);
/* Primitive order after the build (Fbvh::m_bvhpri), nprims entries. */
int sortgpu_accel_prim_order(const sortgpu_accel* accel, uint32_t* out, uint64_t cap);
/* Pre-order (child_cnt, pri_offset, pri_cnt) triples of the tree; returns the number of
 * uint32 values (call with out = NULL to size), or a negative status. */
int64_t sortgpu_accel_tree_dump(const sortgpu_accel* accel, uint32_t* out, uint64_t cap);

/* ---- queries ---- */

/* Closest hit for n rays (Accelerator::GetIntersect, accel/accelerator.h:74; Obvh traversal
 * accel/fast_bvh.hpp:305-504, with Scene::GetIntersect's t = FLT_MAX reset,
 * core/scene.cpp:182-185). d_rays: n records; d_hits: n records. */
int sortgpu_trace_closest(sortgpu_accel* accel, const sortgpu_ray* d_rays, uint32_t n,
                          sortgpu_hit* d_hits, void* stream);

/* sortgpu_trace_closest for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * Same results and traversal order; the order buffers of a large batch live in an 80 MiB array
 * the library holds on each device it runs on (until the process ends) instead of being
 * allocated per call, and the call makes no allocation of any kind. n is at most 4,194,304 rays
 * (larger batches are refused with SORTGPU_ERR_ARG: split them). Every call on one device shares
 * that device's array: calls that overlap on one device (several streams or threads) race and
 * must be serialised by the caller. */
int sortgpu_trace_closest_noalloc(sortgpu_accel* accel, const sortgpu_ray* d_rays, uint32_t n,
                                  sortgpu_hit* d_hits, void* stream);

/* Any hit / occlusion for n rays (Accelerator::IsOccluded, accel/accelerator.h:85; live form:
 * GetIntersect with query_shadow = true, fast_bvh.hpp:362-373, accelerator.cpp:26-39).
 * d_occluded[i] = 1 when some primitive is hit with tmin < t <= tmax, else 0. */
int sortgpu_trace_any(sortgpu_accel* accel, const sortgpu_ray* d_rays, uint32_t n,
                      uint8_t* d_occluded, void* stream);

/* Single-ray compatibility forms for a shim that answers SORT's per-ray calls
 * (Accelerator::GetIntersect, accelerator.h:74, and IsOccluded, accelerator.h:85, called by
 * each worker thread for one ray at a time): host pointers, synchronous, the same traversal and
 * results as the batched calls. Each calling thread gets its own stream and 64 bytes of device
 * staging per device on its first call (released when the thread ends), so worker threads may
 * call concurrently. A wavefront caller uses the batched entry points instead: one round trip per
 * ray costs tens of microseconds. */
int sortgpu_trace_closest1(sortgpu_accel* accel, const sortgpu_ray* ray, sortgpu_hit* hit);
int sortgpu_trace_any1(sortgpu_accel* accel, const sortgpu_ray* ray, int* occluded);

/* What SORT's SurfaceInteraction holds after a closest hit (core/interaction.h), rebuilt on the
 * host from a hit record. */
typedef struct sortgpu_surface {
    float intersect[3]; /* ray(t) = o + t d (math/ray.h:55-57) */
    float t;
    float gnormal[3];   /* triangles: normalize(cross(p2 - p0, p1 - p0)); quads: = normal */
    float normal[3];    /* interpolated shading normal (0 when no normals were given) */
    float tangent[3];   /* interpolated tangent (0 when no tangents were given) */
    float view[3];      /* -d */
    float u, v;         /* interpolated texture coordinates (0 for quads or without texcoords) */
    uint32_t prim;
    uint32_t is_quad;
} sortgpu_surface;

/* Hit reconstruction (setupIntersection, simd/simd_triangle.h:311-341, for triangles; the hit
 * branch of Quad::GetIntersect, shape/quad.cpp:84-91, for quads) from a closest-hit record of
 * `accel` and the ray that produced it, with SORT's float operations in SORT's order. normals and
 * tangents (nverts x 3) and texcoords (nverts x 2) are the mesh's per-vertex attributes in the
 * build's vertex order (SORT's MeshVertex), each optional. Host only; a miss or a primitive id the
 * accelerator does not hold is refused with SORTGPU_ERR_ARG. */
int sortgpu_hit_surface(const sortgpu_accel* accel, const sortgpu_ray* ray, const sortgpu_hit* hit,
                        const float* normals, const float* tangents, const float* texcoords, sortgpu_surface* out);

/* BSSRDF probe query (Accelerator::GetIntersect with BSSRDFIntersections and a material id,
 * accel/fast_bvh.hpp:656-752, simd/simd_triangle.h:412-452, scatteringevent/bssrdf/bssrdf.h:32-56):
 * up to 4 triangle hits per ray whose triangle has material d_ray_material[i]
 * (d_tri_material[prim] per triangle id; quads are never reported). d_hits[4i..4i+3] are the four
 * records in the reference's slot order: filled in the order the walk finds them, then each new
 * hit replaces the first record with the largest t; a hit counts when t is in (tmin, tmax] and
 * below the largest recorded t as it stood before its group of W leaf triangles (that bound is
 * only tightened once a record has been replaced, as in SORT). Unused records are (FLT_MAX,
 * 0xFFFFFFFF, 0, 0); d_counts[i] = records filled (0..4). n_tri_material must cover every
 * triangle (else SORTGPU_ERR_ARG). Stream-ordered. */
int sortgpu_trace_multi(sortgpu_accel* accel, const sortgpu_ray* d_rays, uint32_t n, const uint32_t* d_ray_material,
                        const uint32_t* d_tri_material, uint32_t n_tri_material, sortgpu_hit* d_hits,
                        uint32_t* d_counts, void* stream);

/* Transparent-shadow attenuation for n rays (Scene::GetAttenuation, core/scene.cpp:192-208,
 * looping Accelerator::GetAttenuation, accel/accelerator.cpp:26-68, with the transparent-shadow
 * walk of fast_bvh.hpp:352-405). d_transparency: 4 floats per primitive id (triangles, then quads):
 * the material's transparency r, g, b and 1.0 when the material has transparency, else 0.0 (its
 * transparency then counts as black). Each walk is a closest-hit search that stops, blocked, when
 * a group of W consecutive leaf triangles, or a quad, makes the closest hit on a primitive without
 * transparency (only each group's nearest hit is checked, as in SORT); a blocked ray gets black;
 * a transparent closest hit multiplies the attenuation and the walk goes on from the hit point
 * (o + t d) with (0.001, tmax - t), until a miss or a black attenuation. d_atten[4i..4i+3] = the
 * attenuation r, g, b and the number of surfaces passed (as a float). A walk that passes 4,096
 * surfaces stops there with its attenuation so far and raises SORTGPU_FLAG_ATTEN_CAP.
 * n_transparency = records in d_transparency; fewer than the
 * accelerator's primitives (triangles + quads) is refused (SORTGPU_ERR_ARG). Stream-ordered. */
int sortgpu_trace_attenuation(sortgpu_accel* accel, const sortgpu_ray* d_rays, uint32_t n,
                              const float* d_transparency, uint32_t n_transparency, float* d_atten,
                              void* stream);

/* Traversal work counters for n rays, synchronous: out[0..3] = inner nodes tested, leaves
 * visited, triangles tested, quads tested (the reference's OBVH statistics,
 * fast_bvh.hpp:74-97,407); out[4] = node tests of 8-wide nodes with more than four children (the
 * tests that load all 14 chunks of the record; the others load 8). mode 0 = closest, 1 = any-hit.
 * Results are also written to d_hits / d_occluded exactly as by the plain entry points. Takes the
 * accelerator's error flags afterwards (SORTGPU_ERR_STATE when any is raised; out[] is filled). */
int sortgpu_trace_counters(sortgpu_accel* accel, const sortgpu_ray* d_rays, uint32_t n, int mode,
                           sortgpu_hit* d_hits, uint8_t* d_occluded, uint64_t out[5],
                           void* stream);

/* ---- error flags (no counterpart in SORT, which asserts on the host instead) ----
 * The kernels cannot return a status, so a fault they detect raises a flag on the accelerator
 * (or pool) and the call that launched them still returns SORTGPU_OK. The flags stay raised until
 * they are taken: sortgpu_accel_errors / sortgpu_pool_errors, sortgpu_trace_counters and
 * sortgpu_render_stats take them (read and clear in one atomic step, after the work queued on
 * their stream) and return SORTGPU_ERR_STATE with a message naming each flag when any was raised. */
#define SORTGPU_FLAG_STACK_OVERFLOW 1u /* a traversal stack was full: those rays' results are incomplete */
#define SORTGPU_FLAG_BAD_SCENE 2u      /* sortgpu_pool_trace: a ray named a scene outside the pool (it missed) */
#define SORTGPU_FLAG_ATTEN_CAP 4u      /* sortgpu_trace_attenuation: a walk passed 4,096 surfaces and stopped */

/* Takes the flags raised by the accelerator's queries and renders since the last check, after the
 * work queued on `stream` (synchronises it). *flags (may be NULL) receives them. */
int sortgpu_accel_errors(sortgpu_accel* accel, void* stream, uint32_t* flags);

/* ---- tuning (no counterpart in SORT: the knobs of the persistent traversal kernels) ---- */

/* 0 in any field keeps the default. refill = the compaction threshold of the persistent waves:
 * idle lanes of a wave that trigger refilling them with new rays from the work queue (default 16,
 * 1..64); lds_entries = traversal-stack entries per lane kept in LDS (8 or 16,
 * default 16); blocks_per_cu = cap on resident workgroups per CU (default: occupancy limit);
 * leaf_batch = lanes of a wave holding a leaf that end a node-walking phase (default 32, 1..64);
 * leaf_min = a leaf-testing phase ends when fewer lanes hold a leaf (default 1, 1..64);
 * reorder = traversal order of sortgpu_trace_closest / _any / _attenuation / _multi and
 * sortgpu_pool_trace batches: 0 (default) sorts batches of at least 262,144 rays by origin cell
 * (128^3 grid over the scene box; pooled rays by scene, then by cell in their scene's box) before the
 * launch (stream-ordered device memory of 16 bytes per ray plus sort workspace, from a pool the
 * context owns; a batch sharing one origin, such as camera rays, keeps its order and pays only the
 * sort), 1 = never, 2 = always; shade_order = order in which the path tracer shades each bounce's
 * hits: 0 (default) sorts them by position (a 16^3 grid over the scene box; by scene and an 8^3 grid
 * in multi-request waves) when the wave holds at least 262,144 paths, so shading, the shadow rays and
 * the next bounce's rays run in spatial order, 1 = never, 2 = always. The knobs change speed only:
 * every hit and every image is the same for every setting. */
typedef struct sortgpu_tuning {
    uint32_t refill;
    uint32_t lds_entries;
    uint32_t blocks_per_cu;
    uint32_t leaf_batch;
    uint32_t leaf_min;
    uint32_t reorder;
    uint32_t shade_order;
    uint32_t reserved;
} sortgpu_tuning;

/* Applies to every later query, render and pool trace of the context. */
int sortgpu_ctx_set_tuning(sortgpu_ctx* ctx, const sortgpu_tuning* tuning);

/* ---- batched form: many scenes' rays in one launch over a concatenated tree pool ----
 * (no single counterpart in SORT, which traces one ray against one scene at a time through
 * Accelerator::GetIntersect; each pooled tree is exactly the accelerator it was built as) */

typedef struct sortgpu_pool sortgpu_pool;

/* Copies the records of n accelerators (one width, built on any context, host-only included)
 * into one device allocation, with a device table of their tree views. */
int sortgpu_pool_create(sortgpu_ctx* ctx, sortgpu_accel* const* accels, uint32_t n, sortgpu_pool** out);
void sortgpu_pool_destroy(sortgpu_pool* pool);
/* out[0..3] = scenes, width, device bytes, deepest stack. */
int sortgpu_pool_info(const sortgpu_pool* pool, uint64_t out[4]);
/* One persistent launch over rays of many scenes: d_ray_scene[i] = index of ray i's scene in
 * the pool. any = 0: d_out is n sortgpu_hit; any = 1: n uint8 occlusion flags. */
int sortgpu_pool_trace(sortgpu_pool* pool, const sortgpu_ray* d_rays, const uint32_t* d_ray_scene, uint32_t n,
                       int any, void* d_out, void* stream);
/* Takes the flags raised by the pool's traces since the last check (as sortgpu_accel_errors). */
int sortgpu_pool_errors(sortgpu_pool* pool, void* stream, uint32_t* flags);

/* ---- path tracing (PathTracing::Li, integrator/pathtracing.cpp:37-258) ---- */

#define SORTGPU_MAT_LAMBERT 0    /* Lambert (scatteringevent/bsdf/lambert.cpp)                    */
#define SORTGPU_MAT_GGX 1        /* MicroFacetReflection, GGX, conductor Fresnel (microfacet.cpp)  */
#define SORTGPU_MAT_DIELECTRIC 2 /* rough dielectric, exterior IOR 1 (dielectric.cpp)              */
#define SORTGPU_MAT_DISNEY 3     /* Disney BRDF subset: no transmission / thin / SSS (disney.cpp)  */
#define SORTGPU_MAT_TRANSPARENT 4 /* Transparent (bsdf/transparent.cpp): base = attenuation; light
                                     passes straight through, and a scene holding one answers its
                                     shadow rays with the transparent-shadow walk (GetAttenuation) */

/* One BXDF per material (the configurations' materials; SORT builds them from TSL closures). */
typedef struct sortgpu_material {
    uint32_t type;
    float base[3];     /* Lambert R, GGX R, dielectric reflectance, Disney base color */
    float trans[3];    /* dielectric transmittance */
    float roughness;   /* GGX / dielectric / Disney */
    float ior;         /* dielectric interior IOR */
    float eta[3];      /* conductor Fresnel (GGX) */
    float k[3];
    float metallic, specular, specular_tint, sheen, sheen_tint, clearcoat, clearcoat_gloss;
} sortgpu_material;

/* Shading data over an accelerator's geometry. Every quad of the accelerator is a one-sided
 * area light (AreaLight + Quad, light/area.cpp) with the default white Lambert surface
 * (material/matmanager.h:54-57). */
typedef struct sortgpu_scene_desc {
    const float* normals;          /* nverts x 3 shading normals, or NULL: normalize(cross(p1-p0, p2-p0)) */
    const float* tangents;         /* nverts x 3, or NULL: normalize(p1 - p0) per triangle */
    const uint32_t* tri_material;  /* ntris material indices, or NULL: material 0 */
    const sortgpu_material* materials;
    uint32_t nmaterials;
    const float* quad_radiance;    /* nquads x 3 emitted radiance */
    float env_radiance[3];         /* constant environment (AmbientLight); all 0 = none */
    float clamp;                   /* > 0: each sample's radiance is clamped to [0, clamp] per channel before
                                      it is averaged (the job's clamping, image_evaluation.cpp:222-226);
                                      0 = none; negative or NaN is refused */
} sortgpu_scene_desc;

/* Pinhole perspective camera (PerspectiveCamera, camera/perspective.cpp:26-104). */
typedef struct sortgpu_camera {
    float eye[3];
    float target[3];
    float up[3];
    float fov_y;   /* vertical field of view, radians */
} sortgpu_camera;

typedef struct sortgpu_scene sortgpu_scene;

/* Uploads shading data for `accel` (which must outlive the scene). */
int sortgpu_scene_create(sortgpu_accel* accel, const sortgpu_scene_desc* desc, sortgpu_scene** out);
void sortgpu_scene_destroy(sortgpu_scene* scene);

/* Renders width x height at spp samples per pixel, max_depth bounces (Integrator
 * max_recursive_depth, integrator.h:127), with Russian roulette after the fourth bounce. Writes
 * the per-pixel mean of the valid samples (image_evaluation.cpp:222-244) to d_film_rgb
 * (width * height * 3 floats, device). Rendering of one scene object is serial. */
int sortgpu_render(sortgpu_scene* scene, const sortgpu_camera* camera, uint32_t width, uint32_t height,
                   uint32_t spp, uint32_t max_depth, uint64_t seed, float* d_film_rgb, void* stream);

/* The same frame as sortgpu_render, recording every extension ray PathTracing::li traces
 * (pathtracing.cpp:44-258; bounce 0 = the camera ray) with its closest hit, for the paths of pixels
 * [pix_begin, pix_end): slot ((pixel - pix_begin) * spp + sample) * max_depth + bounce of d_rays,
 * d_hits and d_valid (1 where the path traced that bounce; cleared by the call). Buffers hold
 * (pix_end - pix_begin) * spp * max_depth entries (device). A test and debugging aid: the
 * reference has no such entry point. */
int sortgpu_render_ray_dump(sortgpu_scene* scene, const sortgpu_camera* camera, uint32_t width, uint32_t height,
                            uint32_t spp, uint32_t max_depth, uint64_t seed, uint32_t pix_begin, uint32_t pix_end,
                            sortgpu_ray* d_rays, sortgpu_hit* d_hits, uint8_t* d_valid, void* stream);

/* Batched Integrator::Li (integrator.h:56): radiance along n caller rays. d_keys holds per ray
 * (pixel, sample) — the random-number key of the path. d_rgb: n x 3 floats (device). */
int sortgpu_li(sortgpu_scene* scene, const sortgpu_ray* d_rays, const uint32_t* d_keys, uint32_t n,
               uint32_t max_depth, uint64_t seed, float* d_rgb, void* stream);

/* Totals of the last render / li call (synchronous): out[0] extension rays, out[1] shadow
 * rays, out[2] paths, out[3] microseconds of the last call on its stream (event timing). Takes the
 * accelerator's error flags afterwards (SORTGPU_ERR_STATE when any is raised; out[] is filled). */
int sortgpu_render_stats(sortgpu_scene* scene, uint64_t out[4]);

/* Stage timing of sortgpu_render (no counterpart in SORT; measurement). With on = 1, later
 * renders on the scene bracket each bounce's extension traversal, shading kernel and shadow
 * traversal with timing events on the stream the bounce runs on (the first 1,024 bounces of a
 * render); on = 0 stops it. sortgpu_render_stage_times (synchronous) gives for the last render
 * the summed milliseconds of each stage, out_ms[0] extension traversal, [1] shading, [2] shadow
 * rays, and *bounces the number of bounces timed (0 without timing). Waves on the render's two
 * internal streams overlap, so a stage's time includes what it shared the GPU with. */
int sortgpu_render_set_timing(sortgpu_scene* scene, int on);
int sortgpu_render_stage_times(sortgpu_scene* scene, float out_ms[3], uint32_t* bounces);

/* ---- Served form: continuous-batching render server ------------------------------------------
 * No counterpart in the reference, which renders one job per process (sort.cpp:32-60,
 * core/image_evaluation.cpp). The server keeps scenes, path state and films resident in HBM,
 * admits requests by an HBM budget with least-recently-used scene eviction and interleaves
 * samples of many requests in each wave of the bounce loop. Every request's image equals, bit for
 * bit, what sortgpu_render gives for the same scene, camera, size, spp, depth and seed. */
typedef struct sortgpu_server sortgpu_server;

typedef struct sortgpu_server_opts {
    uint64_t hbm_budget;     /* bytes the server may hold: scenes + films + workspace (0 = 96 GiB) */
    uint32_t wave_paths;     /* path slots per wave (0 = 33,554,432; 248 B of workspace each) */
    uint32_t max_active;     /* requests admitted at once, 1..32 (0 = 32) */
    uint32_t inflight_waves; /* waves in flight at once, 1..4 (0 = 3), each on a stream and path
                                workspace of its own, so their bounce loops overlap */
    uint32_t manual;         /* 1: no worker thread; the caller drives sortgpu_server_step */
    uint32_t host_device;    /* 1: host-only stand-in device (no GPU; films hold a hash of the
                                path key) for tests of the scheduling and failure handling */
    uint32_t fail_alloc_at;  /* test hook: the n-th budgeted allocation (scene load or film)
                                fails as if out of memory (0 = never) */
    uint32_t graphs;         /* 1: replay each wave's bounce loop from a captured HIP graph */
    uint32_t reserved;
} sortgpu_server_opts;

typedef struct sortgpu_request {
    uint32_t scene;          /* id from sortgpu_server_add_scene */
    uint32_t width, height;  /* at most wave_paths pixels */
    uint32_t spp;
    uint32_t max_depth;
    uint32_t reserved;
    uint64_t seed;
    sortgpu_camera camera;
    double timeout_s;        /* > 0: fails with SORTGPU_REQ_TIMED_OUT if not done this long after submission */
} sortgpu_request;

enum {
    SORTGPU_REQ_QUEUED = 0,
    SORTGPU_REQ_ACTIVE = 1,
    SORTGPU_REQ_DONE = 2,
    SORTGPU_REQ_FAILED = 3,
    SORTGPU_REQ_CANCELLED = 4,
    SORTGPU_REQ_TIMED_OUT = 5
};

typedef struct sortgpu_request_info {
    int32_t state;           /* SORTGPU_REQ_* */
    int32_t status;          /* SORTGPU_OK, or the failure's status code */
    uint32_t samples_done;
    uint32_t spp;
    double t_submit, t_admit, t_done;  /* seconds since the server started; 0 = not yet */
} sortgpu_request_info;

int sortgpu_server_create(sortgpu_ctx* ctx, const sortgpu_server_opts* opts, sortgpu_server** out);
/* Cancels what is outstanding, waits for the waves in flight and frees everything. */
void sortgpu_server_destroy(sortgpu_server* srv);
/* Registers a scene: the accelerator's tree (8- or 4-wide; one width per server) and the shading
 * data are copied to the host side of the server, so both may be destroyed afterwards. The scene
 * is uploaded when a request needs it and may be evicted and reloaded later. */
int sortgpu_server_add_scene(sortgpu_server* srv, const sortgpu_accel* accel, const sortgpu_scene_desc* desc,
                             uint32_t* scene_id);
/* Queues a request; invalid parameters are refused here with SORTGPU_ERR_ARG. Thread-safe. */
int sortgpu_server_submit(sortgpu_server* srv, const sortgpu_request* req, uint64_t* id);
int sortgpu_server_cancel(sortgpu_server* srv, uint64_t id);
int sortgpu_server_info(sortgpu_server* srv, uint64_t id, sortgpu_request_info* out);
/* Waits up to timeout_s (< 0: forever) for the request to finish; fills *out either way. */
int sortgpu_server_wait(sortgpu_server* srv, uint64_t id, double timeout_s, sortgpu_request_info* out);
/* Copies a finished request's image (width * height * 3 floats, host memory). */
int sortgpu_server_result(sortgpu_server* srv, uint64_t id, float* rgb, uint64_t n_floats);
/* Copies the failure message of a request (empty when none) into buf (NUL-terminated). */
int sortgpu_server_message(sortgpu_server* srv, uint64_t id, char* buf, uint32_t cap);
/* Forgets a request (cancelling it when outstanding); its id becomes unknown. */
int sortgpu_server_release(sortgpu_server* srv, uint64_t id);
/* Manual mode: one scheduling pass (admit, launch, retire; blocks on the oldest wave in flight).
 * *busy = 1 while requests are queued, active or in flight. */
int sortgpu_server_step(sortgpu_server* srv, int* busy);
/* Waits until nothing is queued, active or in flight (manual mode: steps until then). */
int sortgpu_server_drain(sortgpu_server* srv, double timeout_s);
/* out[0] waves, [1] paths, [2] extension rays, [3] shadow rays, [4] scene loads, [5] evictions,
 * [6] resident scenes, [7] budgeted bytes in use, [8] peak budgeted bytes, [9] done, [10] failed,
 * [11] cancelled, [12] timed out, [13] device busy microseconds, [14] budget, [15] workspace bytes. */
int sortgpu_server_stats(sortgpu_server* srv, uint64_t out[16]);

/* ---- scene files: the .sort stream (host only; no device is touched) ----
 * One file = the global configuration (ImageEvaluation::loadConfig,
 * work/image_evaluation/image_evaluation.cpp:354-378), the material section
 * (MatManager::ParseMatFile, material/matmanager.cpp:50-290, Material::Serialize,
 * material/material.cpp:153-228) and the entity stream (Scene::LoadScene, core/scene.cpp:112-180),
 * as blender-plugin/addons/sortblend/exporter.py writes it. The loader reads every record of the
 * stream the way those readers do and flattens the scene into the arrays the accelerator and the
 * path tracer take. */

#define SORTGPU_ERR_FORMAT (-6) /* unreadable or malformed scene file (message names the offset) */

/* Content of a file that the GPU path does not carry: bits of sortgpu_scene_file_info.skipped;
 * counts[bit] = how many items. */
#define SORTGPU_FILE_PUNCTUAL_LIGHTS 0  /* point, spot and directional lights (light_entity.cpp:25-67) */
#define SORTGPU_FILE_DISK_LIGHTS 1      /* disk area lights (light_entity.cpp:137-148)                  */
#define SORTGPU_FILE_HAIR 2             /* HairVisual curves (entity/visual.cpp:44-100)                */
#define SORTGPU_FILE_VOLUMES 3          /* mesh density grids and volume shaders                       */
#define SORTGPU_FILE_SKY_IMAGE 4        /* HDR sky image (the constant sky colour is kept)             */
#define SORTGPU_FILE_UNMAPPED_MATS 5    /* materials whose shader graph has no GPU BXDF: white Lambert */
#define SORTGPU_FILE_APPROX_MATS 6      /* materials mapped from their constant inputs only (a texture
                                           or other node feeds an input, anisotropic roughness, a
                                           non-GGX distribution, transparency, SSS)                   */
#define SORTGPU_FILE_CAMERA 7           /* lens radius (depth of field) or non-square pixels            */

typedef struct sortgpu_scene_file sortgpu_scene_file;

typedef struct sortgpu_scene_file_info {
    /* global configuration */
    uint32_t threads, spp, image_width, image_height;
    float clamping;
    uint32_t integrator;           /* SID (CRC, core/strid.h:39-60) of the integrator class name   */
    int32_t max_recursive_depth;   /* Integrator::Serialize (integrator.h:85-87)                     */
    int32_t max_bssrdf_bounces;    /* PathTracing only (pathtracing.h:42-45), else 0                 */
    /* camera as stored (PerspectiveCameraEntity, camera_entity.cpp:21-32) */
    float eye[3], up[3], target[3];
    uint32_t camera_width, camera_height;
    float lens_radius, sensor_w, sensor_h;
    int32_t aspect_fit;
    float pixel_aspect_w, pixel_aspect_h, fov;
    /* the pinhole camera the renderer takes: eye, target, up, and the vertical field of view
     * equal to PerspectiveCamera::PreProcess's (camera/perspective.cpp:26-63) for this sensor fit */
    float camera_fov_y;
    uint32_t has_camera;
    /* geometry in primitive-id order: the triangles of every mesh entity in file order (world
     * space, Mesh::ApplyTransform, core/mesh.cpp:26-40), then one quad per square or rectangular
     * area light in file order (AreaLightEntity::Serialize, light_entity.cpp:99-155) */
    uint32_t nverts, ntris, nquads, nmaterials;
    const float* pos;              /* nverts x 3 */
    const float* normals;          /* nverts x 3 (TransformNormal of the normalised normal) */
    const float* uv;               /* nverts x 2: as stored, or Mesh::GenUV's spherical mapping about
                                      the mean vertex for a mesh stored without (core/mesh.cpp:60-74) */
    const float* tangents;         /* nverts x 3: Mesh::GenSmoothTagent (core/mesh.cpp:42-108), the
                                      normalised sum of the tangents of the triangles at each vertex */
    const uint32_t* idx;           /* ntris x 3 */
    const uint32_t* tri_material;  /* ntris: index into materials (a mesh's -1 or unknown id: the last entry) */
    const sortgpu_quad* quads;     /* nquads */
    const float* quad_radiance;    /* nquads x 3 */
    const sortgpu_material* materials; /* the file's materials in order, then SORT's default white Lambert */
    float env_radiance[3];         /* sky colour x intensity; 0.5 when the file has no light at all */
    /* accelerator record after the entities */
    uint32_t accelerator;          /* SID of the accelerator class name */
    int32_t accel_max_depth, accel_max_leaf;
    uint32_t skipped;              /* bits 1 << SORTGPU_FILE_* */
    uint32_t counts[8];
} sortgpu_scene_file_info;

/* Reads a .sort file (or an in-memory copy of one). Every read is bounds-checked: a short,
 * corrupt or unknown-record stream is refused with SORTGPU_ERR_FORMAT and a message naming the
 * byte offset, where SORT would assert. The arrays stay valid until sortgpu_scene_file_destroy. */
int sortgpu_scene_file_load(const char* path, sortgpu_scene_file** out);
int sortgpu_scene_file_parse(const void* data, uint64_t size, sortgpu_scene_file** out);
const sortgpu_scene_file_info* sortgpu_scene_file_get(const sortgpu_scene_file* file);
/* Name of material i (< nmaterials - 1; the last, default material is "default"). */
const char* sortgpu_scene_file_material_name(const sortgpu_scene_file* file, uint32_t i);
void sortgpu_scene_file_destroy(sortgpu_scene_file* file);

/* ---- film output (host + device; no counterpart needed beyond the file and socket formats) ----
 * The film is width x height x 3 floats, row-major from the top row (the layout sortgpu_render
 * writes), in device memory (on_device = 1, converted by a kernel on `stream`; the call
 * synchronises it) or host memory (on_device = 0). */

/* Bytes of the EXR file sortgpu_film_encode_exr writes for a width x height film. */
uint64_t sortgpu_exr_size(uint32_t width, uint32_t height);
/* The OpenEXR file SORT writes for a finished render (Texture2DBase::Output,
 * texture/texturebase.cpp:27-57, through tinyexr's SaveEXR(rgb, w, h, 3, 1)), byte for byte:
 * scanline, channels B G R in half floats (tinyexr's float_to_half_full rounding), no
 * compression. Images below 16 x 16 are refused, as SaveEXR refuses them. out: host memory of at
 * least sortgpu_exr_size bytes. */
int sortgpu_film_encode_exr(const float* film_rgb, int on_device, uint32_t width, uint32_t height, uint8_t* out,
                            uint64_t cap, void* stream);
int sortgpu_film_write_exr(const float* film_rgb, int on_device, uint32_t width, uint32_t height, const char* path,
                           void* stream);

/* Display server packets (DisplayTile / DisplayImageInfo::Process, core/display_mgr.cpp:86-140):
 * BLENDER = the add-on's tile socket (blender-plugin/addons/sortblend/renderer.py:31-86): [bytes
 * after the length, w, h, x, y] + RGBA float32 from the tile's bottom row up; TEV = three
 * UpdateImage packets (R, G, B) of the tile for the tev viewer's socket; TEV_CREATE = its
 * CreateImage packet for the whole image (tile arguments ignored). The film is host memory. Returns
 * the packet size (out = NULL sizes it) or a negative status. */
#define SORTGPU_DISPLAY_BLENDER 0
#define SORTGPU_DISPLAY_TEV 1
#define SORTGPU_DISPLAY_TEV_CREATE 2
int64_t sortgpu_display_packet(int mode, const char* title, const float* film_rgb, uint32_t width, uint32_t height,
                               uint32_t x, uint32_t y, uint32_t tw, uint32_t th, uint8_t* out, uint64_t cap);

typedef struct sortgpu_display sortgpu_display;
/* TCP connection to a display server (DisplayManager::SetupDisplayServer, display_mgr.cpp:33-40). */
int sortgpu_display_connect(const char* host, uint16_t port, sortgpu_display** out);
/* Sends a film as tiles of `tile` pixels (0 = 64, the renderer's tile size) in row-major tile
 * order: BLENDER or TEV mode (TEV sends CreateImage first). */
int sortgpu_display_send_film(sortgpu_display* disp, int mode, const char* title, const float* film_rgb, int on_device,
                              uint32_t width, uint32_t height, uint32_t tile, void* stream);
void sortgpu_display_close(sortgpu_display* disp);

#ifdef __cplusplus
}
#endif

#endif /* SORTGPU_H */
