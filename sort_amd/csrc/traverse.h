// Device-side Fast-BVH traversal: ray setup, slab test, watertight triangle test, quad test
// and the stack walk, one ray per lane.
//
// Arithmetic reproduces SORT's SIMD Obvh/Qbvh operation by operation (the reference computes
// W children / W triangles of one ray per AVX instruction; here each lane computes them for its
// own ray, in the same order):
//   Ray::Prepare                        math/ray.h:26-30,60-69
//   resolveRayData                      simd/simd_ray_utils.h:60-83
//   root slab Intersect(Ray, BBox)      math/bbox.h:124-156
//   IntersectBBox_SIMD                  simd/simd_bbox.h:62-85
//   intersectTriangleInner_SIMD         simd/simd_triangle.h:212-300
//   closest selection                   simd/simd_triangle.h:359-374 (strict t < cur; the lowest
//                                       lane wins ties -> sequential strict update in lane order)
//   Quad::GetIntersect                  shape/quad.cpp:66-94 (accepts t == current, so it wins ties)
//   traversal + push order              accel/fast_bvh.hpp:338-455
//   shadow early exit                   accel/fast_bvh.hpp:362-373
// No FMA: this header is compiled under `#pragma clang fp contract(off)` and -ffp-contract=off.
// Division is IEEE (HIP's default correctly rounded fp32 divide). min/max are written as the
// x86 selects (_mm_min_ps(a,b) = a < b ? a : b), except in the slab test of rays for which no
// slab distance can be NaN: there IEEE min/max (v_min/v_max/v_min3/v_max3) give the same values.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "layout.h"

#pragma clang fp contract(off)

namespace sortgpu {

__device__ __forceinline__ float xmin(float a, float b) { return a < b ? a : b; }   // _mm_min_ps
__device__ __forceinline__ float xmax(float a, float b) { return a > b ? a : b; }   // _mm_max_ps
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; } // std::min
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; } // std::max

__device__ __forceinline__ float sel3(float x, float y, float z, int k) {
    return k == 0 ? x : (k == 1 ? y : z);
}

// Per-ray constants (Ray after Prepare() + Ray8_Data).
struct RayK {
    float ox, oy, oz, tmin, dx, dy, dz, tmax;
    int kx, ky, kz;
    float sx, sy, sz;
    float okx, oky, okz;       // origin permuted into the shear frame
    float rcx, rcy, rcz;       // 1 / clamped dir
    float odx, ody, odz;       // -o / clamped dir
    bool ieee;                 // slab tests may use IEEE min/max (see ieee_slabs_ok)
};

__device__ __forceinline__ void ray_setup(RayK& r) {
    const float ax = fabsf(r.dx), ay = fabsf(r.dy), az = fabsf(r.dz);
    int ky;
    if (ax > ay && ax > az) ky = 0;
    else ky = ay > az ? 1 : 2;
    const int kz = ky == 2 ? 0 : ky + 1;
    const int kx = kz == 2 ? 0 : kz + 1;
    r.kx = kx; r.ky = ky; r.kz = kz;
    const float px = sel3(r.dx, r.dy, r.dz, kx);
    const float py = sel3(r.dx, r.dy, r.dz, ky);
    const float pz = sel3(r.dx, r.dy, r.dz, kz);
    r.sx = -px / py;
    r.sz = -pz / py;
    r.sy = 1.0f / py;
    r.okx = sel3(r.ox, r.oy, r.oz, kx);
    r.oky = sel3(r.ox, r.oy, r.oz, ky);
    r.okz = sel3(r.ox, r.oy, r.oz, kz);
    const float delta = 0.00001f;
    const float cx = fabsf(r.dx) < delta ? (r.dx < 0.0f ? -1.0f : 1.0f) * delta : r.dx;
    const float cy = fabsf(r.dy) < delta ? (r.dy < 0.0f ? -1.0f : 1.0f) * delta : r.dy;
    const float cz = fabsf(r.dz) < delta ? (r.dz < 0.0f ? -1.0f : 1.0f) * delta : r.dz;
    r.rcx = 1.0f / cx;
    r.rcy = 1.0f / cy;
    r.rcz = 1.0f / cz;
    r.odx = -r.ox / cx;
    r.ody = -r.oy / cy;
    r.odz = -r.oz / cz;
}

// Root test against the enlarged scene box; returns tmin or -1.
__device__ __forceinline__ float root_slab(const RayK& r, const float* bb) {
    float tmax = r.tmax, tmin = r.tmin;
    const float o[3] = {r.ox, r.oy, r.oz};
    const float d[3] = {r.dx, r.dy, r.dz};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (d[a] < 0.00001f && d[a] > -0.00001f) {
            if (o[a] > bb[3 + a] || o[a] < bb[a]) return -1.0f;
        } else {
            const float ood = 1.0f / d[a];
            float t1 = (bb[3 + a] - o[a]) * ood;
            float t2 = (bb[a] - o[a]) * ood;
            if (t1 > t2) {
                const float t = t1;
                t1 = t2;
                t2 = t;
            }
            tmin = smax(t1, tmin);
            tmax = smin(t2, tmax);
            if (tmin > tmax) return -1.0f;
        }
    }
    return tmin;
}

// Watertight shear test of one triangle (one lane of intersectTriangleInner_SIMD). Returns true
// with t, u, v when tmin < t <= tmax.
__device__ __forceinline__ bool tri_test(const RayK& r, const Tri& p, float& t_out, float& u_out,
                                         float& v_out) {
    float p0x = sel3(p.p0x, p.p0y, p.p0z, r.kx) - r.okx;
    float p0y = sel3(p.p0x, p.p0y, p.p0z, r.ky) - r.oky;
    float p0z = sel3(p.p0x, p.p0y, p.p0z, r.kz) - r.okz;
    float p1x = sel3(p.p1x, p.p1y, p.p1z, r.kx) - r.okx;
    float p1y = sel3(p.p1x, p.p1y, p.p1z, r.ky) - r.oky;
    float p1z = sel3(p.p1x, p.p1y, p.p1z, r.kz) - r.okz;
    float p2x = sel3(p.p2x, p.p2y, p.p2z, r.kx) - r.okx;
    float p2y = sel3(p.p2x, p.p2y, p.p2z, r.ky) - r.oky;
    float p2z = sel3(p.p2x, p.p2y, p.p2z, r.kz) - r.okz;
    p0x = p0y * r.sx + p0x;
    p0z = p0y * r.sz + p0z;
    p1x = p1y * r.sx + p1x;
    p1z = p1y * r.sz + p1z;
    p2x = p2y * r.sx + p2x;
    p2z = p2y * r.sz + p2z;
    const float e0 = p1x * p2z - p1z * p2x;
    const float e1 = p2x * p0z - p2z * p0x;
    const float e2 = p0x * p1z - p0z * p1x;
    const bool c0 = (e0 >= 0.0f) & (e1 >= 0.0f) & (e2 >= 0.0f);
    const bool c1 = (e0 <= 0.0f) & (e1 <= 0.0f) & (e2 <= 0.0f);
    if (!(c0 | c1)) return false;
    const float det = e0 + (e1 + e2);
    if (!((det < 0.0f) | (det > 0.0f))) return false;  // _CMP_NEQ_OQ: NaN -> false
    const float rcp = 1.0f / det;
    p0y = p0y * r.sy;
    p1y = p1y * r.sy;
    p2y = p2y * r.sy;
    float t = e0 * p0y;
    t = e1 * p1y + t;
    t = e2 * p2y + t;
    t = t * rcp;
    if (!((t > r.tmin) & (t <= r.tmax))) return false;
    t_out = t;
    u_out = e1 * rcp;
    v_out = e2 * rcp;
    return true;
}

// Quad::GetIntersect with limit = current closest t. Returns true with t.
__device__ __forceinline__ bool quad_test(const RayK& r, const QuadRec& q, float limit, float& t_out) {
    const float* m = q.inv;
    float x = r.ox * m[0] + r.oy * m[1] + r.oz * m[2] + m[3];
    float y = r.ox * m[4] + r.oy * m[5] + r.oz * m[6] + m[7];
    float z = r.ox * m[8] + r.oy * m[9] + r.oz * m[10] + m[11];
    const float w = r.ox * m[12] + r.oy * m[13] + r.oz * m[14] + m[15];
    if (!(w == 1.0f)) {
        x = x / w;
        y = y / w;
        z = z / w;
    }
    const float dx = r.dx * m[0] + r.dy * m[1] + r.dz * m[2];
    const float dy = r.dx * m[4] + r.dy * m[5] + r.dz * m[6];
    const float dz = r.dx * m[8] + r.dy * m[9] + r.dz * m[10];
    if (dy == 0.0f) return false;
    const float t = -y / dy;
    if (t > limit || t <= r.tmin || t > r.tmax) return false;
    const float px = x + t * dx;
    const float pz = z + t * dz;
    if (px > q.half_x || px < -q.half_x) return false;
    if (pz > q.half_y || pz < -q.half_y) return false;
    t_out = t;
    return true;
}

// One leaf record: five 16-byte loads.
__device__ __forceinline__ PrimRec load_prim(const PrimRec* __restrict__ base, uint32_t i) {
    const float4* p = reinterpret_cast<const float4*>(base + i);
    const float4 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4];
    PrimRec r;
    r.a.p0x = a.x; r.a.p0y = a.y; r.a.p0z = a.z; r.a.p1x = a.w;
    r.a.p1y = b.x; r.a.p1z = b.y; r.a.p2x = b.z; r.a.p2y = b.w;
    r.a.p2z = c.x; r.b.p0x = c.y; r.b.p0y = c.z; r.b.p0z = c.w;
    r.b.p1x = d.x; r.b.p1y = d.y; r.b.p1z = d.z; r.b.p2x = d.w;
    r.b.p2y = e.x; r.b.p2z = e.y;
    r.id_a = __float_as_uint(e.z);
    r.id_b = __float_as_uint(e.w);
    return r;
}

// Slab test of one W-wide node record. f[j] = entry distance or -1; returns the hit mask.
template <int W>
__device__ __forceinline__ uint32_t node_test(const RayK& r, const NodeRec<W>* __restrict__ nodes,
                                              uint32_t ni, float (&f)[W], uint32_t (&ch)[W]) {
    constexpr int V = W / 4;  // float4 per plane
    const float4* base = reinterpret_cast<const float4*>(nodes + ni);
    float pl[6][W];
#pragma unroll
    for (int v = 0; v < V; v++) {
        const uint4 q = reinterpret_cast<const uint4*>(base)[6 * V + v];
        ch[4 * v + 0] = q.x;
        ch[4 * v + 1] = q.y;
        ch[4 * v + 2] = q.z;
        ch[4 * v + 3] = q.w;
    }
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const float4 q = base[k * V];
        pl[k][0] = q.x;
        pl[k][1] = q.y;
        pl[k][2] = q.z;
        pl[k][3] = q.w;
    }
    if (V == 2) {
        // Children fill the slots in order, so slot 4 empty means slots 4..7 are: their planes
        // are not needed (the REF_EMPTY test discards those slots) and are not loaded. Most
        // 8-wide nodes have at most four children; the loads per lane are what bounds the walk.
        if (ch[4 % W] != REF_EMPTY) {
#pragma unroll
            for (int k = 0; k < 6; k++) {
                const float4 q = base[k * V + (V - 1)];
                pl[k][4 % W] = q.x;
                pl[k][5 % W] = q.y;
                pl[k][6 % W] = q.z;
                pl[k][7 % W] = q.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 6; k++) pl[k][4 % W] = pl[k][5 % W] = pl[k][6 % W] = pl[k][7 % W] = 0.0f;
        }
    }
    uint32_t mask = 0;
    if (r.ieee) {
#pragma unroll
        for (int j = 0; j < W; j++) {
            const float x1 = r.odx + r.rcx * pl[1][j], x2 = r.odx + r.rcx * pl[0][j];
            const float y1 = r.ody + r.rcy * pl[3][j], y2 = r.ody + r.rcy * pl[2][j];
            const float z1 = r.odz + r.rcz * pl[5][j], z2 = r.odz + r.rcz * pl[4][j];
            const float fmn = fmaxf(fmaxf(fmaxf(r.tmin, fminf(x1, x2)), fminf(y1, y2)), fminf(z1, z2));
            const float fmx = fminf(fminf(fminf(r.tmax, fmaxf(x1, x2)), fmaxf(y1, y2)), fmaxf(z1, z2));
            const bool hit = (ch[j] != REF_EMPTY) & (fmn <= fmx);
            f[j] = hit ? fmn : -1.0f;
            mask |= hit ? (1u << j) : 0u;
        }
    } else {
        // Opaque copies: the compiler must not share this rare path's slab distances with the
        // path above (hoisting them ahead of the branch keeps all 6*W of them live at once).
        float odx = r.odx, ody = r.ody, odz = r.odz, rcx = r.rcx, rcy = r.rcy, rcz = r.rcz;
        asm volatile("" : "+v"(odx), "+v"(ody), "+v"(odz), "+v"(rcx), "+v"(rcy), "+v"(rcz));
#pragma unroll
        for (int j = 0; j < W; j++) {
            float fmn = r.tmin, fmx = r.tmax;
            float t1 = odx + rcx * pl[1][j];
            float t2 = odx + rcx * pl[0][j];
            fmn = xmax(fmn, xmin(t1, t2));
            fmx = xmin(fmx, xmax(t1, t2));
            t1 = ody + rcy * pl[3][j];
            t2 = ody + rcy * pl[2][j];
            fmn = xmax(fmn, xmin(t1, t2));
            fmx = xmin(fmx, xmax(t1, t2));
            t1 = odz + rcz * pl[5][j];
            t2 = odz + rcz * pl[4][j];
            fmn = xmax(fmn, xmin(t1, t2));
            fmx = xmin(fmx, xmax(t1, t2));
            const bool hit = (ch[j] != REF_EMPTY) & (fmn <= fmx);
            f[j] = hit ? fmn : -1.0f;
            mask |= hit ? (1u << j) : 0u;
        }
    }
    return mask;
}

// Whether IEEE min/max may stand in for the x86 selects in this ray's slab tests against a tree
// with root box `bb`. The two differ only when an operand is NaN (and in the sign of a zero
// result, which only ever enters comparisons). With a finite origin and direction, the clamped
// direction d' has |d'| >= 1e-5, so 1/d' is finite (|1/d'| <= 1e5) and -o/d' is finite or
// infinite but not NaN; with every box coordinate p within kIeeeBoxBound, (1/d') * p is finite,
// so -o/d' + (1/d') * p is never inf - inf and no slab distance is NaN; a window with no NaN adds
// none. Empty child slots may produce anything: the REF_EMPTY test discards their result.
// Any bound below FLT_MAX / 1e5 is sound; 1e12 is far beyond every float32 scene (at 5e12 the
// triangle test's t numerator already overflows) and lets a tree with real hits sit above it.
constexpr float kIeeeBoxBound = 1e12f;
__device__ __forceinline__ bool ieee_slabs_ok(const RayK& r, const float* bb) {
    bool ok = __builtin_isfinite(r.ox) & __builtin_isfinite(r.oy) & __builtin_isfinite(r.oz) &
              __builtin_isfinite(r.dx) & __builtin_isfinite(r.dy) & __builtin_isfinite(r.dz) &
              !__builtin_isnan(r.tmin) & !__builtin_isnan(r.tmax);
#pragma unroll
    for (int k = 0; k < 6; k++) ok &= fabsf(bb[k]) <= kIeeeBoxBound;
    return ok;
}

}  // namespace sortgpu
