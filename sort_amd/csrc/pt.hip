// Wavefront path tracer: SORT's PathTracing::Li bounce loop (integrator/pathtracing.cpp:44-258)
// with EvaluateDirect's light-sampled + BSDF-sampled next-event estimation and power-2 MIS
// (integrator/integratormethod.cpp:28-30,100-192), restructured as stages over fixed path slots:
//
//   raygen   camera rays (PerspectiveCamera::GenerateRay, camera/perspective.cpp:66-104) or caller rays
//   extend   closest-hit traversal of the live queue (k_trace, device-side count)
//   shade    emission at the first vertex, NEE (up to two shadow rays per vertex, written to
//            per-path slots), BSDF continuation, Russian roulette, compaction of survivors by
//            wave ballot + one atomic per wave
//   shadow   any-hit traversal of the compacted shadow queue; an occluded ray clears its
//            contribution slot from inside the traversal (with transparent materials: the
//            attenuation walk, then resolve scales or clears each slot)
//   finalize fold the pending contributions, average each pixel's valid samples in sample order
//            (image_evaluation.cpp:222-244: NaN/inf samples dropped from the mean)
//
// Path results do not depend on queue order: all randomness is keyed by (seed, pixel, sample,
// bounce, call) and every accumulation into a path happens in a fixed order.
#include <hip/hip_runtime.h>

#include "pt_launch.h"
#include "shading.h"
#include "trace_launch.h"

namespace sortgpu {

constexpr int kPtBlock = 256;

__device__ __forceinline__ float mis(float f, float g) { return (f * f) / (f * f + g * g); }

// Distribution1D::SampleDiscrete (core/samplemethod.h:158-176).
__device__ __forceinline__ int sample_discrete(const float* cdf, uint32_t count, float u, float* pdf) {
    uint32_t lo = 0, hi = count + 1;  // lower_bound over cdf[0..count]
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    uint32_t off = (u <= 0.0f) ? 0u : lo - 1u;
    if (off == 0)
        while (off < count && cdf[off + 1] == 0.0f) off++;
    if (off == count) {
        *pdf = 0.0f;
        return 0;
    }
    *pdf = cdf[off + 1] - cdf[off];
    return (int)off;
}

// Wave-aggregated append: every lane with `pred` gets a unique slot in [0, *counter).
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t* counter) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return 0;
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

__device__ __forceinline__ void put_ray(float4* q, uint32_t i, V3 o, float tmin, V3 d, float tmax) {
    q[2 * i] = make_float4(o.x, o.y, o.z, tmin);
    q[2 * i + 1] = make_float4(d.x, d.y, d.z, tmax);
}

// ---- raygen ------------------------------------------------------------------------------------
__device__ __forceinline__ int chunk_of_path(const ChunkTable& ct, uint32_t p) {
    int k = 0;
    for (int i = 1; i < (int)ct.n; i++)
        if (ct.c[i].path_base <= p) k = i;
    return k;
}

// A new path's record: throughput 1, nothing gathered, its pixel, sample and request.
__device__ __forceinline__ void init_path(PathState ps, uint32_t p, uint32_t pix, uint32_t s, uint32_t req) {
    float4* r = ps.rec + 4 * (size_t)p;
    r[0] = make_float4(1.f, 1.f, 1.f, __uint_as_float(pix));
    r[1] = make_float4(0.f, 0.f, 0.f, __uint_as_float(s));
    r[2] = make_float4(0.f, 0.f, 0.f, __uint_as_float(req));
    r[3] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Camera rays of every chunk (PerspectiveCamera::GenerateRay in closed form: raster -> camera
// through the inverse of Scale(W,H,1) Scale(.5,-.5,1) Translate(1,-1,0) Perspective(xs, ys),
// then the orthonormal ViewLookat basis).
__global__ __launch_bounds__(kPtBlock) void k_raygen(const WaveDesc* __restrict__ w, PathState ps, Queues q) {
    const ChunkTable& ct = w->ct;
    const uint32_t p = blockIdx.x * kPtBlock + threadIdx.x;
    if (p == 0) q.count[0] = ct.total_paths;
    if (p >= ct.total_paths) return;
    const ChunkDev& c = ct.c[chunk_of_path(ct, p)];
    const ReqDev& R = w->reqs[c.req];
    const uint32_t local = p - c.path_base;
    const uint32_t pix = local % c.npix;
    const uint32_t s = c.sample0 + local / c.npix;
    const CamDev& cam = R.cam;
    const uint32_t x = pix % cam.width, y = pix / cam.width;
    const Rng4 j = rng4(R.seed, pix, s, kCameraBounce, 0);
    const float rx = (float)x + j.a, ry = (float)y + j.b;
    const float cx = 2.0f * rx / (float)cam.width - 1.0f;
    const float cy = 1.0f - 2.0f * ry / (float)cam.height;
    const V3 dc = normalize(v3(cx / cam.x_scale, cy / cam.y_scale, 1.0f));
    const V3 d = cam.right * dc.x + cam.up * dc.y + cam.fwd * dc.z;
    put_ray(q.ray[0], p, cam.eye, 0.0f, d, FLT_MAX);
    q.pid[0][p] = p;
    q.rsc[0][p] = R.scene;
    init_path(ps, p, pix, s, c.req);
}

// Caller-provided rays (the batched form of Integrator::Li) of request 0: keys[2i] = pixel,
// keys[2i+1] = sample.
__global__ __launch_bounds__(kPtBlock) void k_raygen_rays(const float4* rays, const uint32_t* keys, uint32_t n,
                                                          const ReqDev* __restrict__ reqs, PathState ps, Queues q) {
    const uint32_t p = blockIdx.x * kPtBlock + threadIdx.x;
    if (p == 0) q.count[0] = n;
    if (p >= n) return;
    q.ray[0][2 * p] = rays[2 * p];
    q.ray[0][2 * p + 1] = rays[2 * p + 1];
    q.pid[0][p] = p;
    q.rsc[0][p] = reqs[0].scene;
    init_path(ps, p, keys[2 * p], keys[2 * p + 1], 0u);
}

__global__ void k_zero_counts(uint32_t* count, int which) {
    if (threadIdx.x == 0) {
        count[which] = 0;
        count[2] = 0;
    }
}

// Records the extension rays of the current queue and their hits into the ray dump (paths of
// pixels in the dump's range only).
__global__ __launch_bounds__(kPtBlock) void k_dump(uint32_t bounce, int cur, PathState ps, Queues q, RayDumpDev d) {
    const uint32_t n = q.count[cur];
    for (uint32_t i = blockIdx.x * kPtBlock + threadIdx.x; i < n; i += gridDim.x * kPtBlock) {
        const uint32_t p = q.pid[cur][i];
        const float4* pr = ps.rec + 4 * (size_t)p;
        const uint32_t pix = __float_as_uint(pr[0].w), s = __float_as_uint(pr[1].w);
        if (pix < d.pix_begin || pix >= d.pix_end || s >= d.spp || bounce >= d.depth) continue;
        const size_t k = ((size_t)(pix - d.pix_begin) * d.spp + s) * d.depth + bounce;
        d.rays[2 * k] = q.ray[cur][2 * i];
        d.rays[2 * k + 1] = q.ray[cur][2 * i + 1];
        d.hits[k] = q.hit[i];
        d.valid[k] = 1;
    }
}

__global__ void k_add_totals(const uint32_t* count, int ext_slot, unsigned long long* totals) {
    if (threadIdx.x == 0) {
        totals[0] += count[ext_slot];
        totals[1] += count[2];
    }
}

// Hit reconstruction at a triangle (setupIntersection, simd/simd_triangle.h:311-341): shading
// normal and tangent interpolated with w = 1 - u - v.
struct Surf {
    V3 p, n, t;
    uint32_t mat;
    int32_t light;  // light index when the hit primitive is an area light
};

__device__ __forceinline__ V3 load3(const float* a, uint32_t i) { return v3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }

__device__ __forceinline__ Surf surface(const SceneDev& sc, V3 o, V3 d, float t, uint32_t prim, float u, float v) {
    Surf s;
    s.p = o + t * d;
    s.light = -1;
    if (prim < sc.ntris) {
        const uint32_t i0 = sc.idx[3 * prim], i1 = sc.idx[3 * prim + 1], i2 = sc.idx[3 * prim + 2];
        const float w = 1.0f - u - v;
        if (sc.nrm) {
            s.n = normalize(w * load3(sc.nrm, i0) + u * load3(sc.nrm, i1) + v * load3(sc.nrm, i2));
        } else {
            const V3 p0 = load3(sc.pos, i0), p1 = load3(sc.pos, i1), p2 = load3(sc.pos, i2);
            s.n = normalize(cross(p1 - p0, p2 - p0));
        }
        if (sc.tan) {
            s.t = normalize(w * load3(sc.tan, i0) + u * load3(sc.tan, i1) + v * load3(sc.tan, i2));
        } else {
            s.t = normalize(load3(sc.pos, i1) - load3(sc.pos, i0));
        }
        s.mat = sc.tri_mat ? sc.tri_mat[prim] : 0u;
    } else {
        const uint32_t qi = prim - sc.ntris;
        const int32_t li = sc.quad_light[qi];
        const LightRec& l = sc.lights[li];
        s.n = quad_normal(l);                      // TransformNormal(DIR_UP)
        s.t = v3(l.mat[2], l.mat[6], l.mat[10]);  // TransformVector((0,0,1))
        s.mat = sc.nmats;                          // default material: white Lambert
        s.light = li;
    }
    return s;
}

// Shading frame (ScatteringEvent ctor, scatteringevent.cpp:44-49).
struct Frame {
    V3 t, n, b;
};
__device__ __forceinline__ Frame make_frame(V3 n_in, V3 tangent) {
    Frame f;
    f.n = normalize(n_in);
    f.b = normalize(cross(f.n, tangent));
    f.t = normalize(cross(f.b, f.n));
    return f;
}
__device__ __forceinline__ V3 to_local(const Frame& f, V3 v) { return v3(dot(v, f.t), dot(v, f.n), dot(v, f.b)); }
__device__ __forceinline__ V3 to_world(const Frame& f, V3 v) {
    return v3(v.x * f.t.x + v.y * f.n.x + v.z * f.b.x, v.x * f.t.y + v.y * f.n.y + v.z * f.b.y,
              v.x * f.t.z + v.y * f.n.z + v.z * f.b.z);
}

// ---- spatial order of a bounce's hits ----------------------------------------------------------
// Hits of a bounce land anywhere in the scene. Shading them in the order of their position makes
// the lanes of a wave gather nearby vertices and materials, and, because shading appends shadow
// rays and continuations in its own order, the shadow rays and the next bounce's rays leave from
// nearby points too, so the two traversals that follow walk shared subtrees. A count-bounded bucket
// sort (the queue length is known only on the device): per-slice LDS histograms, a scan per bucket
// over the slices, a scan over the buckets, and a scatter. Paths are keyed, so no result depends on
// the order.
__device__ __forceinline__ uint32_t spread3_4(uint32_t x) {  // 4 bits to every third position
    x &= 0xfu;
    x = (x | (x << 4)) & 0x0c3u;
    x = (x | (x << 2)) & 0x249u;
    return x;
}
__device__ __forceinline__ uint32_t grid_cell(float t, float cells) {
    return (uint32_t)fminf(fmaxf(t * cells, 0.0f), cells - 1.0f);  // NaN -> 0
}

struct SortBox {
    float lo[3], inv[3];
};

__device__ __forceinline__ uint32_t hit_bucket(const Queues& q, int cur, uint32_t i, const TreeView* trees,
                                               const SortBox& box) {
    const float4 h = q.hit[i];
    if (__float_as_uint(h.y) == 0xFFFFFFFFu) return kSortBuckets - 1;  // misses: shade reads only the sky
    const float4 r0 = q.ray[cur][2 * (size_t)i], r1 = q.ray[cur][2 * (size_t)i + 1];
    const float p[3] = {r0.x + h.x * r1.x, r0.y + h.x * r1.y, r0.z + h.x * r1.z};
    if (!trees) {  // 16^3 cells over the scene box, Morton order
        uint32_t c[3];
        for (int k = 0; k < 3; k++) c[k] = grid_cell((p[k] - box.lo[k]) * box.inv[k], 16.0f);
        return spread3_4(c[0]) | (spread3_4(c[1]) << 1) | (spread3_4(c[2]) << 2);
    }
    // multi-scene waves: 8 scene classes (scene index mod 8) x 8^3 cells of that scene's box
    const uint32_t sid = q.rsc[cur][i];
    const float* b = trees[sid].bbox;
    uint32_t m = 0;
    for (int k = 0; k < 3; k++) {
        const float w = b[3 + k] - b[k];
        const uint32_t c = grid_cell(w > 0.0f ? (p[k] - b[k]) / w : 0.0f, 8.0f);
        m |= spread3_4(c) << k;
    }
    return ((sid & 7u) << 9) | (m & 0x1ffu);
}

__global__ __launch_bounds__(256) void k_sort_count(int cur, const TreeView* trees, SortBox box, Queues q) {
    __shared__ uint32_t h[kSortBuckets];
    for (uint32_t j = threadIdx.x; j < kSortBuckets; j += 256) h[j] = 0;
    __syncthreads();
    const uint32_t n = q.count[cur];
    const uint32_t chunk = (n + kSortBlocks - 1) / kSortBlocks;
    const uint32_t lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += 256) {
        const uint32_t key = hit_bucket(q, cur, i, trees, box);
        q.skey[i] = key;
        atomicAdd(&h[key], 1u);
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < kSortBuckets; j += 256) q.shist[blockIdx.x * kSortBuckets + j] = h[j];
}

// Per bucket: exclusive offsets of the slices within the bucket, and the bucket's total. A
// workgroup takes 16 buckets; each of its lanes sums 32 consecutive slices of one bucket (the 16
// lanes of a slice row read 64 contiguous bytes), the 16 partial sums of a bucket are scanned in
// LDS, and each lane writes the running offsets of its slices.
constexpr uint32_t kScanBuckets = 16, kScanGroups = 256 / kScanBuckets, kScanRun = kSortBlocks / kScanGroups;
static_assert(kSortBuckets % kScanBuckets == 0 && kSortBlocks % kScanGroups == 0, "scan tiling");
__global__ __launch_bounds__(256) void k_sort_scan_slices(Queues q) {
    __shared__ uint32_t part[kScanGroups][kScanBuckets];
    const uint32_t jb = threadIdx.x % kScanBuckets, g = threadIdx.x / kScanBuckets;
    const uint32_t j = blockIdx.x * kScanBuckets + jb;
    uint32_t* col = q.shist + (size_t)g * kScanRun * kSortBuckets + j;
    uint32_t v[kScanRun], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanRun; k++) v[k] = col[(size_t)k * kSortBuckets];
#pragma unroll
    for (uint32_t k = 0; k < kScanRun; k++) sum += v[k];
    part[g][jb] = sum;
    __syncthreads();
    uint32_t run = 0;
    for (uint32_t k = 0; k < g; k++) run += part[k][jb];
    if (g == kScanGroups - 1) q.sbase[j] = run + sum;
#pragma unroll
    for (uint32_t k = 0; k < kScanRun; k++) {
        col[(size_t)k * kSortBuckets] = run;
        run += v[k];
    }
}

// Exclusive scan of the bucket totals (one workgroup of 1024 lanes, 4 buckets each): a scan
// within each 64-lane wave by lane shuffles, then over the 16 wave totals.
__global__ __launch_bounds__(1024) void k_sort_scan_buckets(Queues q) {
    __shared__ uint32_t wsum[16];
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    uint32_t v[4], sum = 0;
    for (int k = 0; k < 4; k++) {
        v[k] = q.sbase[4 * t + k];
        sum += v[k];
    }
    uint32_t incl = sum;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t x = __shfl_up(incl, d);
        if (lane >= d) incl += x;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t run = incl - sum;
    for (uint32_t k = 0; k < w; k++) run += wsum[k];
    for (int k = 0; k < 4; k++) {
        q.sbase[4 * t + k] = run;
        run += v[k];
    }
}

__global__ __launch_bounds__(256) void k_sort_scatter(int cur, Queues q) {
    __shared__ uint32_t off[kSortBuckets];
    for (uint32_t j = threadIdx.x; j < kSortBuckets; j += 256)
        off[j] = q.sbase[j] + q.shist[blockIdx.x * kSortBuckets + j];
    __syncthreads();
    const uint32_t n = q.count[cur];
    const uint32_t chunk = (n + kSortBlocks - 1) / kSortBlocks;
    const uint32_t lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += 256) q.sperm[atomicAdd(&off[q.skey[i]], 1u)] = i;
}

// ---- shade -------------------------------------------------------------------------------------
// `order` (or null): the queue entries in the order to shade them.
__global__ __launch_bounds__(kPtBlock) void k_shade(const SceneDev* __restrict__ scenes,
                                                    const ReqDev* __restrict__ reqs, uint32_t bounce, int cur,
                                                    PathState ps, Queues q, const uint32_t* __restrict__ order) {
    const uint32_t n = q.count[cur];
    const int nxt = cur ^ 1;
    for (uint32_t base = blockIdx.x * kPtBlock; base < n; base += gridDim.x * kPtBlock) {
        const uint32_t k = base + threadIdx.x;
        const uint32_t i = (order && k < n) ? order[k] : k;
        bool cont = false, sh0 = false, sh1 = false;
        V3 ro = v3(0, 0, 0), rd = v3(0, 0, 0), so = v3(0, 0, 0), s0d = v3(0, 0, 0), s1d = v3(0, 0, 0);
        float s0max = 0.f, s1max = 0.f;
        uint32_t p = 0, scene = 0;
        if (i < n) {
            p = q.pid[cur][i];
            float4* pr = ps.rec + 4 * (size_t)p;  // the path's record: one 64-byte read, one write
            const float4 Tp = pr[0], Lp = pr[1], pc0 = pr[2], pc1 = pr[3];
            const uint32_t req = __float_as_uint(pc0.w);
            const ReqDev& R = reqs[req];
            scene = R.scene;
            const SceneDev& sc = scenes[scene];
            const uint32_t seed = R.seed, max_depth = R.max_depth;
            const float4 h = q.hit[i];
            const float4 r0 = q.ray[cur][2 * i], r1 = q.ray[cur][2 * i + 1];
            const V3 o = v3(r0.x, r0.y, r0.z), d = v3(r1.x, r1.y, r1.z);
            V3 L = v3(Lp.x, Lp.y, Lp.z);
            // fold the previous vertex's resolved NEE contributions
            L = L + v3(pc0.x, pc0.y, pc0.z) + v3(pc1.x, pc1.y, pc1.z);
            float4 c0 = make_float4(0.f, 0.f, 0.f, 0.f), c1 = c0;
            V3 T = v3(Tp.x, Tp.y, Tp.z);
            const uint32_t prim = __float_as_uint(h.y);
            if (prim == 0xFFFFFFFFu) {
                if (bounce == 0 && sc.env_light >= 0) L = L + sc.env_L;  // Scene::Le on a camera miss
            } else {
                const uint32_t key = __float_as_uint(Tp.w), smp = __float_as_uint(Lp.w);
                const Surf s = surface(sc, o, d, h.x, prim, h.z, h.w);
                if (bounce == 0 && s.light >= 0) {  // SurfaceInteraction::Le -> AreaLight::Le
                    const LightRec& l = sc.lights[s.light];
                    const float c = saturate(dot(-d, s.n));
                    if (c != 0.0f) L = L + l.L;
                }
                const MatRec m = sc.mats[s.mat];
                const Frame fr = make_frame(s.n, s.t);
                const V3 wo_w = -d;
                const V3 wo = to_local(fr, wo_w);
                const Rng4 ra = rng4(seed, key, smp, bounce, 0);
                const Rng4 rb = rng4(seed, key, smp, bounce, 1);
                const Rng4 rc = rng4(seed, key, smp, bounce, 2);
                const Rng4 rf_nee = rng4(seed, key, smp, bounce, 3);
                const Rng4 rf_cont = rng4(seed, key, smp, bounce, 4);
                // ---- next-event estimation (EvaluateDirect) ----
                float pick_pdf = 0.0f;
                const int li = sc.nlights ? sample_discrete(sc.light_cdf, sc.nlights, ra.a, &pick_pdf) : 0;
                if (pick_pdf > 0.0f) {
                    const LightRec& l = sc.lights[li];
                    V3 wi, Li = v3(0, 0, 0);
                    float light_pdf = 0.0f, smax = 0.0f;
                    if (l.kind == 0) {  // AreaLight::sample_l -> Quad::Sample_l
                        const float uu = 2.0f * ra.b - 1.0f, vv = 2.0f * ra.c - 1.0f;
                        const V3 lp = xform_point3(l.mat, v3(l.half_x * uu, 0.0f, l.half_y * vv));
                        const V3 nl = xform_vec3(l.mat, v3(0.0f, 1.0f, 0.0f));
                        const V3 delta = lp - s.p;
                        wi = normalize(delta);
                        const float dd = dot(-wi, nl);
                        light_pdf = dd <= 0.0f ? 0.0f : dot(delta, delta) / (l.area * dd);
                        const float len = length(delta);
                        smax = len - 0.01f;
                        Li = l.L;
                    } else {  // AmbientLight / SkyLight::sample_l
                        wi = uniform_sample_sphere(ra.b, ra.c);
                        light_pdf = kInvFourPi;
                        smax = FLT_MAX;
                        Li = l.L;
                    }
                    if (light_pdf > 0.0f && !is_black(Li)) {
                        const V3 wil = to_local(fr, wi);
                        const V3 f = mat_f(m, wo, wil);
                        if (!is_black(f)) {
                            const float bp = mat_pdf(m, wo, wil);
                            const float w = mis(light_pdf, bp);
                            const V3 c = T * (Li * f * w / light_pdf) / pick_pdf;
                            c0 = make_float4(c.x, c.y, c.z, 0.f);
                            sh0 = true;
                            so = s.p;
                            s0d = wi;
                            s0max = smax;
                        }
                    }
                    // BSDF-sampled half with a separate BsdfSample (integratormethod.cpp:150-189)
                    BsdfSampleOut bs = mat_sample(m, wo, rb.b, rb.c, rf_nee);
                    if (bs.pdf == 0.0f) bs.f = v3(0, 0, 0);
                    if (!is_black(bs.f) && bs.pdf != 0.0f) {
                        const V3 wi2 = to_world(fr, bs.wi);
                        float lp2 = 0.0f, tl = -1.0f;
                        V3 Li2 = v3(0, 0, 0);
                        if (l.kind == 0) {  // Shape::Pdf + AreaLight::Le(ray)
                            tl = quad_hit(l, s.p, wi2);
                            if (tl >= 0.0f) {
                                const V3 nl = quad_normal(l);
                                const V3 hp = s.p + tl * wi2;
                                const V3 delta = s.p - hp;
                                const float dn = saturate(dot(normalize(delta), nl));
                                lp2 = dn <= 0.0f ? 0.0f : dot(delta, delta) / (l.area * dn);
                                if (saturate(dot(-wi2, nl)) != 0.0f) Li2 = l.L;
                            }
                        } else {
                            lp2 = kInvFourPi;
                            tl = FLT_MAX;
                            Li2 = l.L;
                        }
                        if (lp2 > 0.0f && tl >= 0.0f && !is_black(Li2)) {
                            const float w2 = mis(bs.pdf, lp2);
                            const V3 c = T * (Li2 * bs.f * w2 / bs.pdf) / pick_pdf;
                            c1 = make_float4(c.x, c.y, c.z, 0.f);
                            sh1 = true;
                            s1d = wi2;
                            s1max = tl - 0.001f;
                            so = s.p;
                        }
                    }
                }
                // ---- continuation (pathtracing.cpp:179-249) ----
                BsdfSampleOut cs = mat_sample(m, wo, rc.b, rc.c, rf_cont);
                if (cs.pdf == 0.0f) cs.f = v3(0, 0, 0);
                if (!is_black(cs.f) && cs.pdf != 0.0f) {
                    T = T * (cs.f / cs.pdf);
                    if (luminance(T) != 0.0f) {
                        cont = true;
                        if (bounce > 3 && maxc(T) < 0.1f) {
                            const float q_rr = fmaxf(0.05f, 1.0f - maxc(T));
                            if (rc.d < q_rr) cont = false;
                            else T = T / (1.0f - q_rr);
                        }
                        if (bounce + 1 >= max_depth) cont = false;
                        ro = s.p;
                        rd = to_world(fr, cs.wi);
                    }
                }
            }
            pr[0] = make_float4(T.x, T.y, T.z, Tp.w);
            pr[1] = make_float4(L.x, L.y, L.z, Lp.w);
            pr[2] = make_float4(c0.x, c0.y, c0.z, pc0.w);
            pr[3] = c1;
        }
        // compaction: survivors and shadow rays append with one atomic per wave
        const uint32_t ci = wave_append(cont, &q.count[nxt]);
        if (cont) {
            put_ray(q.ray[nxt], ci, ro, 0.0001f, rd, FLT_MAX);
            q.pid[nxt][ci] = p;
            q.rsc[nxt][ci] = scene;
        }
        const uint32_t si0 = wave_append(sh0, &q.count[2]);
        if (sh0) {
            put_ray(q.sray, si0, so, 0.01f, s0d, s0max);
            q.ssid[si0] = 2 * p;
            q.ssc[si0] = scene;
        }
        const uint32_t si1 = wave_append(sh1, &q.count[2]);
        if (sh1) {
            put_ray(q.sray, si1, so, 0.001f, s1d, s1max);
            q.ssid[si1] = 2 * p + 1;
            q.ssc[si1] = scene;
        }
    }
}

// ---- resolve -----------------------------------------------------------------------------------
// Transparent shadows: every shadow ray's attenuation (r, g, b, surfaces) was written over the
// first half of its ray record; a black one clears its contribution slot, any other scales it
// (EvaluateDirect's attenuation factor, integratormethod.cpp:134-142 and 181-185; applied here to
// the finished contribution instead of first in the product, which moves float rounding only).
__global__ __launch_bounds__(kPtBlock) void k_resolve_att(PathState ps, Queues q) {
    const uint32_t n = q.count[2];
    for (uint32_t i = blockIdx.x * kPtBlock + threadIdx.x; i < n; i += gridDim.x * kPtBlock) {
        const float4 a = q.sray[2 * (size_t)i];
        float4* c = ps.rec + path_slot(q.ssid[i]);
        const float4 v = *c;
        *c = (a.x == 0.0f && a.y == 0.0f && a.z == 0.0f) ? make_float4(0.f, 0.f, 0.f, v.w)
                                                         : make_float4(v.x * a.x, v.y * a.y, v.z * a.z, v.w);
    }
}

// ---- finalize ----------------------------------------------------------------------------------
// One thread per (chunk, pixel): adds the chunk's samples of that pixel, in sample order, to the
// request's film (a request has at most one chunk per wave, so no two threads share a pixel).
__global__ __launch_bounds__(kPtBlock) void k_finalize(const WaveDesc* __restrict__ w, PathState ps) {
    const ChunkTable& ct = w->ct;
    const ReqDev* reqs = w->reqs;
    const uint32_t g = blockIdx.x * kPtBlock + threadIdx.x;
    if (g >= ct.total_pix) return;
    int k = 0;
    for (int i = 1; i < (int)ct.n; i++)
        if (ct.c[i].pix_base <= g) k = i;
    const ChunkDev& c = ct.c[k];
    const uint32_t pix = g - c.pix_base;
    float4* film = reqs[c.req].film;
    const float cl = reqs[c.req].clamp;
    float4 acc = film[pix];
    for (uint32_t s = 0; s < c.nsamp; s++) {
        const uint32_t p = c.path_base + s * c.npix + pix;
        const float4* pr = ps.rec + 4 * (size_t)p;
        const float4 l = pr[1], a = pr[2], b = pr[3];
        float r = l.x + a.x + b.x, gg = l.y + a.y + b.y, bb = l.z + a.z + b.z;
        if (cl > 0.0f) {  // Spectrum::Clamp(0, clamping): clamp() of math/utils.h:89-93, NaN kept
            r = r > cl ? cl : r;
            r = r < 0.0f ? 0.0f : r;
            gg = gg > cl ? cl : gg;
            gg = gg < 0.0f ? 0.0f : gg;
            bb = bb > cl ? cl : bb;
            bb = bb < 0.0f ? 0.0f : bb;
        }
        if (isfinite(r) && isfinite(gg) && isfinite(bb)) {
            acc.x += r;
            acc.y += gg;
            acc.z += bb;
            acc.w += 1.0f;
        }
    }
    film[pix] = acc;
}

// Per-ray radiance for the caller-ray form.
__global__ __launch_bounds__(kPtBlock) void k_finalize_rays(PathState ps, uint32_t n, float* out) {
    const uint32_t p = blockIdx.x * kPtBlock + threadIdx.x;
    if (p >= n) return;
    const float4* pr = ps.rec + 4 * (size_t)p;
    const float4 l = pr[1], a = pr[2], b = pr[3];
    out[3 * p] = l.x + a.x + b.x;
    out[3 * p + 1] = l.y + a.y + b.y;
    out[3 * p + 2] = l.z + a.z + b.z;
}

__global__ __launch_bounds__(kPtBlock) void k_resolve_film(const float4* film, uint32_t npix, float* rgb) {
    const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
    if (i >= npix) return;
    const float4 a = film[i];
    const float inv = a.w > 0.0f ? 1.0f / a.w : 0.0f;
    rgb[3 * i] = a.x * inv;
    rgb[3 * i + 1] = a.y * inv;
    rgb[3 * i + 2] = a.z * inv;
}

__global__ void k_zero_f4(float4* a, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        a[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---- host driver -------------------------------------------------------------------------------
static inline unsigned blocks_for(uint64_t n) { return (unsigned)((n + kPtBlock - 1) / kPtBlock); }

hipError_t pt_zero_f4(float4* a, uint32_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    unsigned b = blocks_for(n);
    if (b > 4096) b = 4096;
    hipLaunchKernelGGL(k_zero_f4, dim3(b), dim3(kPtBlock), 0, s, a, n);
    return hipGetLastError();
}

// Runs the bounce loop for the paths already in queue 0 (count q.count[0] <= npaths).
hipError_t pt_bounces(const PtLaunch& L, uint32_t npaths, hipStream_t s) {
    const unsigned shade_blocks = std::min<unsigned>(blocks_for(npaths), (unsigned)L.num_cus * 8u);
    hipError_t e;
    for (uint32_t b = 0; b < L.max_depth; b++) {
        const int cur = (int)(b & 1u), nxt = cur ^ 1;
        hipEvent_t* tev = nullptr;  // this bounce's six stage events, when timed
        if (L.stage_ev && *L.stage_used + 6 <= L.stage_cap) {
            tev = L.stage_ev + *L.stage_used;
            *L.stage_used += 6;
        }
        auto mark = [&](int k) -> hipError_t { return tev ? hipEventRecord(tev[k], s) : hipSuccess; };
        TraceArgs ta = L.trace;
        ta.rays = L.q.ray[cur];
        ta.n = npaths;
        ta.n_dev = &L.q.count[cur];
        ta.out = L.q.hit;
        if (L.trees) {
            ta.trees = L.trees;
            ta.ray_scene = L.q.rsc[cur];
            ta.n_trees = L.n_trees;
            ta.width = L.tree_width;
            ta.max_stack = L.tree_max_stack;
        }
        if ((e = mark(0)) != hipSuccess || (e = launch_trace(ta, false, false, s, L.num_cus)) != hipSuccess ||
            (e = mark(1)) != hipSuccess)
            return e;
        if (L.dump.rays) hipLaunchKernelGGL(k_dump, dim3(shade_blocks), dim3(kPtBlock), 0, s, b, cur, L.ps, L.q, L.dump);
        hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, s, L.q.count, nxt);
        const uint32_t* order = nullptr;
        if (L.sort_min && npaths >= L.sort_min && L.q.sperm) {
            SortBox box;
            for (int k = 0; k < 3; k++) {
                const float w = L.trace.tree.bbox[3 + k] - L.trace.tree.bbox[k];
                box.lo[k] = L.trace.tree.bbox[k];
                box.inv[k] = w > 0.0f ? 1.0f / w : 0.0f;
            }
            hipLaunchKernelGGL(k_sort_count, dim3(kSortBlocks), dim3(256), 0, s, cur, L.trees, box, L.q);
            hipLaunchKernelGGL(k_sort_scan_slices, dim3(kSortBuckets / kScanBuckets), dim3(256), 0, s, L.q);
            hipLaunchKernelGGL(k_sort_scan_buckets, dim3(1), dim3(1024), 0, s, L.q);
            hipLaunchKernelGGL(k_sort_scatter, dim3(kSortBlocks), dim3(256), 0, s, cur, L.q);
            order = L.q.sperm;
        }
        if ((e = mark(2)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_shade, dim3(shade_blocks), dim3(kPtBlock), 0, s, L.scenes, L.wave->reqs, b, cur, L.ps,
                           L.q, order);
        if ((e = hipGetLastError()) != hipSuccess || (e = mark(3)) != hipSuccess || (e = mark(4)) != hipSuccess)
            return e;
        TraceArgs sa = ta;
        sa.rays = L.q.sray;
        sa.n = 2 * npaths;
        sa.n_dev = &L.q.count[2];
        if (L.trees ? L.transps != nullptr : L.transp != nullptr) {
            // Visibility::GetAttenuation: each walk's result overwrites its own ray's first half
            // (a lane reads its ray before it writes the result, and no other lane reads it)
            sa.out = L.q.sray;
            sa.out_stride = 2;
            if (L.trees) {
                sa.ray_scene = L.q.ssc;
                sa.transps = L.transps;
            } else {
                sa.transp = L.transp;
            }
            if ((e = launch_attenuation(sa, s, L.num_cus)) != hipSuccess) return e;
            hipLaunchKernelGGL(k_resolve_att, dim3(shade_blocks), dim3(kPtBlock), 0, s, L.ps, L.q);
        } else {
            // occluded shadow rays clear their contribution slots inside the any-hit launch
            sa.out = nullptr;
            sa.occ_slot = L.q.ssid;
            sa.occ_rec = L.ps.rec;
            if (L.trees) sa.ray_scene = L.q.ssc;
            if ((e = launch_trace(sa, true, false, s, L.num_cus)) != hipSuccess) return e;
        }
        if ((e = mark(5)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_add_totals, dim3(1), dim3(64), 0, s, L.q.count, cur, L.q.totals);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t pt_wave_paths(const PtLaunch& L, hipStream_t s) {
    if (L.paths_cap == 0) return hipSuccess;
    hipLaunchKernelGGL(k_raygen, dim3(blocks_for(L.paths_cap)), dim3(kPtBlock), 0, s, L.wave, L.ps, L.q);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return pt_bounces(L, L.paths_cap, s);
}

hipError_t pt_wave_film(const PtLaunch& L, hipStream_t s) {
    if (L.pix_cap == 0) return hipSuccess;
    hipLaunchKernelGGL(k_finalize, dim3(blocks_for(L.pix_cap)), dim3(kPtBlock), 0, s, L.wave, L.ps);
    return hipGetLastError();
}

hipError_t pt_wave(const PtLaunch& L, hipStream_t s) {
    hipError_t e = pt_wave_paths(L, s);
    return e != hipSuccess ? e : pt_wave_film(L, s);
}

hipError_t pt_resolve_film(const float4* film, uint32_t npix, float* rgb, hipStream_t s) {
    hipLaunchKernelGGL(k_resolve_film, dim3(blocks_for(npix)), dim3(kPtBlock), 0, s, film, npix, rgb);
    return hipGetLastError();
}

hipError_t pt_li_rays(const PtLaunch& L, const float4* rays, const uint32_t* keys, uint32_t n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_raygen_rays, dim3(blocks_for(n)), dim3(kPtBlock), 0, s, rays, keys, n, L.wave->reqs, L.ps,
                       L.q);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = pt_bounces(L, n, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_finalize_rays, dim3(blocks_for(n)), dim3(kPtBlock), 0, s, L.ps, n, out);
    return hipGetLastError();
}

}  // namespace sortgpu
