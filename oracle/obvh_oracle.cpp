// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of SORT's Fast-BVH (OBVH, 8-wide, AVX; QBVH, 4-wide, SSE) build and
// ray queries, used by the test-suite as the parity checker for the HIP product path and by
// bench.py as the measured CPU baseline ("port"). Nothing in the product (sort_amd/) links,
// loads or calls this file.
//
// Arithmetic follows the reference operation by operation: no FMA (mul then add, the
// reference's simd_mad_ps, simd/simd_intel.h:262-264), IEEE division for every reciprocal
// (simd_rcp_ps, simd/simd_intel.h:259-261), x86 min/max operand order, ordered compares.
// Build with: -O3 -mavx -mno-fma -ffp-contract=off (see oracle/Makefile).
//
// Function-by-function citations (paths relative to SORT's src/):
//   BBox::Union / HalfSurfaceArea / MaxAxisId / Delta   math/bbox.cpp:85-127
//   Intersect(Ray, BBox) (scalar root slab)             math/bbox.h:124-156
//   Ray::Prepare, majorAxis                             math/ray.h:26-30,60-69
//   resolveRayData (Ray8_Data / Ray4_Data)              simd/simd_ray_utils.h:60-83
//   IntersectBBox_SIMD                                  simd/simd_bbox.h:62-85
//   intersectTriangleInner_SIMD / intersectTriangle_SIMD simd/simd_triangle.h:212-300,350-383
//   Triangle::GetIntersect (scalar, double fallback)    shape/triangle.cpp:25-114
//   Triangle::GetBBox                                   shape/triangle.cpp:116-136
//   Quad::GetIntersect / GetBBox                        shape/quad.cpp:66-108
//   Matrix::TransformPoint / TransformVector            math/matrix.cpp:70-92
//   Scene bbox (+0.1% per side)                         core/scene.cpp:145-154
//   Bvh_Primitive::SetPrimitive (centroid)              accel/bvh_utils.h:35-38
//   sah / pickBestSplit                                 accel/bvh_utils.h:56-116
//   Fbvh::Build / splitNode / makeLeaf                  accel/fast_bvh.hpp:109-270
//   calcBoundingBox / calcBoundingBoxSIMD               accel/fast_bvh.hpp:100-107,273-302
//   Fbvh::GetIntersect (closest + query_shadow mode)    accel/fast_bvh.hpp:305-504
//   Accelerator::GetAttenuation (opaque)                accel/accelerator.cpp:26-39
//
// How each part is pinned (tests/test_oracle_*.py): closed forms for ray setup, the slab test,
// the triangle and quad tests; brute force over every primitive for closest-hit t (bitwise)
// and any-hit (bitwise); BVH invariants (prim coverage, box containment, leaf size, SAH
// termination); the SIMD wrapper known answers of unittests/simd.hpp (rcp == 1/x bitwise,
// min-reduction, movemask).

#include <immintrin.h>
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <queue>
#include <thread>
#include <utility>
#include <vector>

namespace orc {

struct V3 {
    float x = 0.f, y = 0.f, z = 0.f;
    V3() = default;
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float  operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
static inline V3 operator-(const V3& a, const V3& b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 operator+(const V3& a, const V3& b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 operator*(const V3& a, float s) { return V3(a.x * s, a.y * s, a.z * s); }

// std::min / std::max exactly as libstdc++ defines them (operand order matters for -0/+0).
static inline float smin(float a, float b) { return (b < a) ? b : a; }
static inline float smax(float a, float b) { return (a < b) ? b : a; }

// math/bbox.h:24-102, math/bbox.cpp
struct BBox {
    V3 mn{FLT_MAX, FLT_MAX, FLT_MAX};
    V3 mx{-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void Union(const V3& p) {  // bbox.cpp:106-113
        for (int i = 0; i < 3; i++) {
            if (p[i] < mn[i]) mn[i] = p[i];
            if (p[i] > mx[i]) mx[i] = p[i];
        }
    }
    void Union(const BBox& b) {  // bbox.cpp:115-120
        for (int i = 0; i < 3; i++) {
            mn[i] = smin(mn[i], b.mn[i]);
            mx[i] = smax(mx[i], b.mx[i]);
        }
    }
    float HalfSurfaceArea() const {  // bbox.cpp:85-88
        V3 o = mx - mn;
        return o.x * o.y + o.y * o.z + o.z * o.x;
    }
    unsigned MaxAxisId() const {  // bbox.cpp:95-104
        V3 o = mx - mn;
        if (o.x > o.y && o.x > o.z) return 0;
        if (o.y > o.x && o.y > o.z) return 1;
        return 2;
    }
    float Delta(unsigned k) const { return mx[k] - mn[k]; }
};
static inline BBox UnionF(const BBox& a, const BBox& b) {  // bbox.h:115-122
    BBox r;
    for (int i = 0; i < 3; i++) {
        r.mn[i] = smin(a.mn[i], b.mn[i]);
        r.mx[i] = smax(a.mx[i], b.mx[i]);
    }
    return r;
}

// math/ray.h:37-93 (only the fields the intersection path reads)
struct Ray {
    V3 o, d;
    float fmin = 0.f, fmax = FLT_MAX;
    int kx = 0, ky = 0, kz = 0;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    void Prepare() {  // ray.h:60-69 with majorAxis ray.h:26-30
        const float ax = std::fabs(d[0]), ay = std::fabs(d[1]), az = std::fabs(d[2]);
        if (ax > ay && ax > az) ky = 0;
        else ky = ay > az ? 1 : 2;
        kz = (ky + 1) % 3;
        kx = (kz + 1) % 3;
        const float px = d[kx], py = d[ky], pz = d[kz];
        sx = -px / py;
        sz = -pz / py;
        sy = 1.0f / py;
    }
};

// math/matrix.cpp:70-92 ; quad transforms
struct Mat4 {
    float m[16];
    V3 TransformPoint(const V3& p) const {
        const float x = p.x * m[0] + p.y * m[1] + p.z * m[2] + m[3];
        const float y = p.x * m[4] + p.y * m[5] + p.z * m[6] + m[7];
        const float z = p.x * m[8] + p.y * m[9] + p.z * m[10] + m[11];
        const float w = p.x * m[12] + p.y * m[13] + p.z * m[14] + m[15];
        if (w == 1.0f) return V3(x, y, z);
        return V3(x / w, y / w, z / w);  // Vector3 / scalar, math/vector3.h:218-220
    }
    V3 TransformVector(const V3& v) const {
        return V3(v.x * m[0] + v.y * m[1] + v.z * m[2], v.x * m[4] + v.y * m[5] + v.z * m[6],
                  v.x * m[8] + v.y * m[9] + v.z * m[10]);
    }
};

struct Quad {
    Mat4 mat, inv;
    float sizeX, sizeY;
    BBox bbox;
    void ComputeBBox() {  // quad.cpp:96-108
        const float hx = sizeX * 0.5f, hy = sizeY * 0.5f;
        bbox = BBox();
        bbox.Union(mat.TransformPoint(V3(hx, 0.f, hy)));
        bbox.Union(mat.TransformPoint(V3(hx, 0.f, -hy)));
        bbox.Union(mat.TransformPoint(V3(-hx, 0.f, hy)));
        bbox.Union(mat.TransformPoint(V3(-hx, 0.f, -hy)));
    }
    // quad.cpp:66-94. `limit` is intersect->t. Returns true and writes t on a hit.
    bool Intersect(const Ray& r, float limit, float* t_out) const {
        const V3 o = inv.TransformPoint(r.o);
        const V3 d = inv.TransformVector(r.d);
        if (d.y == 0.0f) return false;
        const float t = -o.y / d.y;
        if (t > limit || t <= r.fmin || t > r.fmax) return false;
        const V3 p(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
        const float hx = sizeX * 0.5f, hy = sizeY * 0.5f;
        if (p.x > hx || p.x < -hx) return false;
        if (p.z > hy || p.z < -hy) return false;
        *t_out = t;
        return true;
    }
};

struct Hit {
    float t;
    uint32_t prim;
    float u, v;
};

struct Stats {
    uint64_t rays = 0, nodes = 0, leaves = 0, tris = 0, quads = 0;
};

// ------------------------------------------------------------------------------------------
// SIMD wrappers with the reference's semantics (simd/simd_intel.h).
template <int W> struct Simd;

template <> struct Simd<8> {  // simd_intel.h:175-313
    using T = __m256;
    static T set1(float f) { return _mm256_set1_ps(f); }
    static T load(const float* p) { return _mm256_load_ps(p); }
    static T add(T a, T b) { return _mm256_add_ps(a, b); }
    static T sub(T a, T b) { return _mm256_sub_ps(a, b); }
    static T mul(T a, T b) { return _mm256_mul_ps(a, b); }
    static T rcp(T a) { return _mm256_div_ps(_mm256_set1_ps(1.0f), a); }
    static T mad(T a, T b, T c) { return _mm256_add_ps(_mm256_mul_ps(a, b), c); }
    static T pick(T m, T a, T b) { return _mm256_blendv_ps(b, a, m); }
    static T ceq(T a, T b) { return _mm256_cmp_ps(a, b, _CMP_EQ_OQ); }
    static T cneq(T a, T b) { return _mm256_cmp_ps(a, b, _CMP_NEQ_OQ); }
    static T cle(T a, T b) { return _mm256_cmp_ps(a, b, _CMP_LE_OQ); }
    static T clt(T a, T b) { return _mm256_cmp_ps(a, b, _CMP_LT_OQ); }
    static T cge(T a, T b) { return _mm256_cmp_ps(a, b, _CMP_GE_OQ); }
    static T cgt(T a, T b) { return _mm256_cmp_ps(a, b, _CMP_GT_OQ); }
    static T and_(T a, T b) { return _mm256_and_ps(a, b); }
    static T or_(T a, T b) { return _mm256_or_ps(a, b); }
    static int movemask(T a) { return _mm256_movemask_ps(a); }
    static T min(T a, T b) { return _mm256_min_ps(a, b); }
    static T max(T a, T b) { return _mm256_max_ps(a, b); }
    static float minreduce(T s) {  // simd_intel.h:302-311
        const __m256i m0 = _mm256_set_epi32(1, 0, 3, 2, 5, 4, 7, 6);
        const __m256i m1 = _mm256_set_epi32(2, 3, 0, 1, 6, 7, 4, 5);
        __m256 a = s;
        __m256 sh = _mm256_permutevar_ps(a, m1);
        a = _mm256_min_ps(sh, a);
        sh = _mm256_permutevar_ps(a, m0);
        __m256 r = _mm256_min_ps(sh, a);
        alignas(32) float f[8];
        _mm256_store_ps(f, r);
        return f[0] < f[4] ? f[0] : f[4];
    }
    static void store(float* p, T a) { _mm256_store_ps(p, a); }
};

template <> struct Simd<4> {  // simd_intel.h:34-173
    using T = __m128;
    static T set1(float f) { return _mm_set1_ps(f); }
    static T load(const float* p) { return _mm_load_ps(p); }
    static T add(T a, T b) { return _mm_add_ps(a, b); }
    static T sub(T a, T b) { return _mm_sub_ps(a, b); }
    static T mul(T a, T b) { return _mm_mul_ps(a, b); }
    static T rcp(T a) { return _mm_div_ps(_mm_set1_ps(1.0f), a); }
    static T mad(T a, T b, T c) { return _mm_add_ps(_mm_mul_ps(a, b), c); }
    static T pick(T m, T a, T b) { return _mm_blendv_ps(b, a, m); }
    static T ceq(T a, T b) { return _mm_cmpeq_ps(a, b); }
    // simd_intel.h:137-138: the SSE wrapper's != is unordered (true for NaN), unlike the AVX
    // wrapper's _CMP_NEQ_OQ; its one use (det != 0) never sees a NaN det.
    static T cneq(T a, T b) { return _mm_cmpneq_ps(a, b); }
    static T cle(T a, T b) { return _mm_cmple_ps(a, b); }
    static T clt(T a, T b) { return _mm_cmplt_ps(a, b); }
    static T cge(T a, T b) { return _mm_cmpge_ps(a, b); }
    static T cgt(T a, T b) { return _mm_cmpgt_ps(a, b); }
    static T and_(T a, T b) { return _mm_and_ps(a, b); }
    static T or_(T a, T b) { return _mm_or_ps(a, b); }
    static int movemask(T a) { return _mm_movemask_ps(a); }
    static T min(T a, T b) { return _mm_min_ps(a, b); }
    static T max(T a, T b) { return _mm_max_ps(a, b); }
    static float minreduce(T s) {  // simd_intel.h:134-137
        const __m128 t = _mm_min_ps(s, _mm_shuffle_ps(s, s, _MM_SHUFFLE(2, 3, 0, 1)));
        const __m128 r = _mm_min_ps(t, _mm_shuffle_ps(t, t, _MM_SHUFFLE(1, 0, 3, 2)));
        return _mm_cvtss_f32(r);
    }
    static void store(float* p, T a) { _mm_store_ps(p, a); }
};

static inline float mask_true() {
    const uint32_t u = 0xffffffffu;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// simd_ray_utils.h:42-83
template <int W> struct alignas(32) RayW {
    using S = Simd<W>;
    typename S::T ori_dir[3], rcp_dir[3], ori[3], scale[3];
};
template <int W> static inline void resolveRayData(const Ray& ray, RayW<W>& s) {
    using S = Simd<W>;
    constexpr float delta = 0.00001f;
    float dd[3];
    for (int i = 0; i < 3; i++) {
        const float di = ray.d[i];
        dd[i] = std::fabs(di) < delta ? (di < 0.0f ? -1.0f : 1.0f) * delta : di;
    }
    for (int i = 0; i < 3; i++) {
        s.rcp_dir[i] = S::set1(1.0f / dd[i]);
        s.ori_dir[i] = S::set1(-ray.o[i] / dd[i]);
        s.ori[i] = S::set1(ray.o[i]);
    }
    s.scale[0] = S::set1(ray.sx);
    s.scale[1] = S::set1(ray.sy);
    s.scale[2] = S::set1(ray.sz);
}

// math/bbox.h:124-156 (scalar root slab)
static inline float IntersectScalar(const Ray& ray, const BBox& bb) {
    float tmax = ray.fmax, tmin = ray.fmin;
    for (int a = 0; a < 3; a++) {
        if (ray.d[a] < 0.00001f && ray.d[a] > -0.00001f) {
            if (ray.o[a] > bb.mx[a] || ray.o[a] < bb.mn[a]) return -1.0f;
        } else {
            const float ood = 1.0f / ray.d[a];
            float t1 = (bb.mx[a] - ray.o[a]) * ood;
            float t2 = (bb.mn[a] - ray.o[a]) * ood;
            if (t1 > t2) std::swap(t1, t2);
            tmin = smax(t1, tmin);
            tmax = smin(t2, tmax);
            if (tmin > tmax) return -1.0f;
        }
    }
    return tmin;
}

// ------------------------------------------------------------------------------------------
// Tree (fast_bvh.h:55-83) and packed leaf records (simd_triangle.h:56-199, simd_bbox.h:47-58)
template <int W> struct alignas(32) BBoxW {
    float mn[3][W];
    float mx[3][W];
    float mask[W];
};
template <int W> struct alignas(32) TriW {
    float p[9][W];  // p0x p0y p0z p1x p1y p1z p2x p2y p2z
    float mask[W];
    uint32_t prim[W];
    uint32_t count;
};
template <int W> struct Node {
    BBoxW<W> bbox;
    std::vector<TriW<W>> tris;
    std::vector<uint32_t> others;  // quad indices (prim id = ntris + q)
    std::unique_ptr<Node> children[W];
    unsigned pri_cnt = 0, pri_offset = 0, child_cnt = 0;
    Node(unsigned off, unsigned cnt) : pri_cnt(cnt), pri_offset(off) {}
};

struct Prim {  // bvh_utils.h:28-46 ; prim id stands in for the Primitive pointer
    uint32_t id;
    V3 centroid;
};

struct Scene {
    const float* pos = nullptr;
    const uint32_t* idx = nullptr;
    uint32_t ntris = 0;
    std::vector<Quad> quads;
    std::vector<BBox> pbox;  // per primitive
    BBox bbox;               // enlarged scene bbox
};

// BSSRDFIntersections (scatteringevent/bssrdf/bssrdf.h:32-56): at most four records, maxt only
// resolved when a record is replaced.
struct MultiHits {
    Hit rec[4];
    uint32_t cnt = 0;
    float maxt = FLT_MAX;
};

template <int W> struct Bvh {
    Scene sc;
    std::vector<Prim> prims;
    std::unique_ptr<Node<W>> root;
    unsigned maxLeaf = 16, maxDepth = 28, depth = 0;
    bool valid = false;
    uint64_t nodeCount = 0, leafCount = 0;

    V3 vtx(uint32_t t, int k) const {
        const uint32_t vi = sc.idx[3 * t + k];
        return V3(sc.pos[3 * vi], sc.pos[3 * vi + 1], sc.pos[3 * vi + 2]);
    }
    const BBox& GetBBox(uint32_t id) const { return sc.pbox[id]; }

    // bvh_utils.h:56-58
    static float sah(unsigned left, unsigned right, const BBox& lb, const BBox& rb, const BBox& b) {
        return (left * lb.HalfSurfaceArea() + right * rb.HalfSurfaceArea()) / b.HalfSurfaceArea();
    }

    // bvh_utils.h:69-116
    float pickBestSplit(unsigned& axis, float& splitPos, const BBox& node_bbox, unsigned start,
                        unsigned end) const {
        constexpr unsigned N = 16;
        constexpr float INV = 1.0f / (float)N;
        BBox inner;
        for (unsigned i = start; i < end; i++) inner.Union(prims[i].centroid);
        const unsigned primitive_num = end - start;
        axis = inner.MaxAxisId();
        float min_sah = FLT_MAX;
        unsigned bin[N];
        BBox bbox[N];
        BBox rbox[N - 1];
        std::memset(bin, 0, sizeof(bin));
        const float split_start = inner.mn[axis];
        const float split_delta = inner.Delta(axis) * INV;
        if (split_delta == 0.0f) return FLT_MAX;
        const float inv_split_delta = 1.0f / split_delta;
        for (unsigned i = start; i < end; i++) {
            int index = (int)((prims[i].centroid[axis] - split_start) * inv_split_delta);
            index = std::min(index, (int)(N - 1));
            ++bin[index];
            bbox[index].Union(GetBBox(prims[i].id));
        }
        rbox[N - 2].Union(bbox[N - 1]);
        for (int i = N - 3; i >= 0; i--) rbox[i] = UnionF(rbox[i + 1], bbox[i + 1]);
        unsigned left = bin[0];
        BBox lbox = bbox[0];
        float pos = split_delta + split_start;
        for (unsigned i = 0; i < N - 1; i++) {
            const float s = sah(left, primitive_num - left, lbox, rbox[i], node_bbox);
            if (s < min_sah) {
                min_sah = s;
                splitPos = pos;
            }
            left += bin[i + 1];
            lbox.Union(bbox[i + 1]);
            pos += split_delta;
        }
        return min_sah;
    }

    BBox calcBoundingBox(const Node<W>* n) const {  // fast_bvh.hpp:100-107
        BBox b;
        if (!n) return b;
        for (unsigned i = n->pri_offset; i < n->pri_offset + n->pri_cnt; i++) b.Union(GetBBox(prims[i].id));
        return b;
    }

    void makeLeaf(Node<W>* node, unsigned start, unsigned end, unsigned d) {  // fast_bvh.hpp:213-270
        node->pri_cnt = end - start;
        node->pri_offset = start;
        node->child_cnt = 0;
        depth = std::max(depth, d);
        TriW<W> g;
        unsigned cnt = 0;
        auto flush = [&]() {
            if (!cnt) return;
            for (unsigned l = 0; l < (unsigned)W; l++) g.mask[l] = l < cnt ? mask_true() : 0.0f;
            g.count = cnt;
            node->tris.push_back(g);
            cnt = 0;
        };
        for (unsigned i = start; i < end; i++) {
            const uint32_t id = prims[i].id;
            if (id < sc.ntris) {
                for (int k = 0; k < 3; k++) {
                    const V3 p = vtx(id, k);
                    g.p[3 * k + 0][cnt] = p.x;
                    g.p[3 * k + 1][cnt] = p.y;
                    g.p[3 * k + 2][cnt] = p.z;
                }
                g.prim[cnt] = id;
                if (++cnt == (unsigned)W) flush();
            } else {
                node->others.push_back(id - sc.ntris);
            }
        }
        if (cnt) {  // partial group: unused lanes masked off (simd_triangle.h:139-145)
            for (unsigned l = cnt; l < (unsigned)W; l++) {
                for (int c = 0; c < 9; c++) g.p[c][l] = 0.0f;
                g.prim[l] = 0xffffffffu;
            }
            flush();
        }
        leafCount++;
    }

    void calcBoundingBoxSIMD(Node<W>* node) {  // fast_bvh.hpp:273-302
        for (int i = 0; i < W; i++) {
            const BBox bb = calcBoundingBox(node->children[i].get());
            for (int a = 0; a < 3; a++) {
                node->bbox.mn[a][i] = bb.mn[a];
                node->bbox.mx[a][i] = bb.mx[a];
            }
            node->bbox.mask[i] = node->children[i] ? mask_true() : 0.0f;
        }
    }

    void splitNode(Node<W>* node, const BBox& node_bbox, unsigned d) {  // fast_bvh.hpp:138-211
        const unsigned start = node->pri_offset;
        const unsigned end = start + node->pri_cnt;
        if (node->pri_cnt <= maxLeaf || d == maxDepth) {
            makeLeaf(node, start, end, d);
            return;
        }
        std::queue<std::pair<unsigned, unsigned>> to_split, done;
        to_split.push(std::make_pair(start, end));
        while (!to_split.empty() && to_split.size() + done.size() < (unsigned)W) {
            const auto cur = to_split.front();
            to_split.pop();
            const unsigned s = cur.first, e = cur.second, cnt = e - s;
            unsigned axis;
            float pos;
            const float sahv = pickBestSplit(axis, pos, node_bbox, s, e);
            if (sahv >= cnt || cnt <= maxLeaf) {
                done.push(cur);
            } else {
                const auto cmp = [pos, axis](const Prim& p) { return p.centroid[axis] < pos; };
                const auto middle = std::partition(prims.data() + s, prims.data() + e, cmp);
                const unsigned mid = (unsigned)(middle - prims.data());
                if (mid == s || mid == e) {
                    done.push(cur);
                } else {
                    to_split.push(std::make_pair(s, mid));
                    to_split.push(std::make_pair(mid, e));
                }
            }
        }
        if (to_split.size() + done.size() == 1) {
            makeLeaf(node, start, end, d);
            return;
        }
        auto populate = [&](std::queue<std::pair<unsigned, unsigned>>& q) {
            while (!q.empty()) {
                const auto c = q.front();
                q.pop();
                node->children[node->child_cnt++].reset(new Node<W>(c.first, c.second - c.first));
            }
        };
        populate(to_split);
        populate(done);
        for (unsigned j = 0; j < node->child_cnt; ++j) {
            const BBox bb = calcBoundingBox(node->children[j].get());
            splitNode(node->children[j].get(), bb, d + 1);
        }
        calcBoundingBoxSIMD(node);
        nodeCount += node->child_cnt;
    }

    void Build() {  // fast_bvh.hpp:109-136 ; scene bbox scene.cpp:145-154
        const uint32_t n = sc.ntris + (uint32_t)sc.quads.size();
        sc.pbox.resize(n);
        for (uint32_t t = 0; t < sc.ntris; t++) {  // triangle.cpp:116-136
            BBox b;
            b.Union(vtx(t, 0));
            b.Union(vtx(t, 1));
            b.Union(vtx(t, 2));
            sc.pbox[t] = b;
        }
        for (uint32_t q = 0; q < sc.quads.size(); q++) {
            sc.quads[q].ComputeBBox();
            sc.pbox[sc.ntris + q] = sc.quads[q].bbox;
        }
        BBox sb;
        for (uint32_t i = 0; i < n; i++) sb.Union(sc.pbox[i]);
        const V3 delta = (sb.mx - sb.mn) * 0.001f;
        sb.mn = sb.mn - delta;
        sb.mx = sb.mx + delta;
        sc.bbox = sb;
        if (!n) return;
        prims.resize(n);
        for (uint32_t i = 0; i < n; i++) {
            prims[i].id = i;
            const BBox& b = sc.pbox[i];
            prims[i].centroid = (b.mx + b.mn) * 0.5f;
        }
        root.reset(new Node<W>(0, n));
        splitNode(root.get(), sc.bbox, 1u);
        valid = true;
        nodeCount++;
    }

    // simd_bbox.h:62-85
    static inline int IntersectBBox(const Ray& ray, const RayW<W>& sr, const BBoxW<W>& bb,
                                    typename Simd<W>::T& f_min) {
        using S = Simd<W>;
        f_min = S::set1(ray.fmin);
        typename S::T f_max = S::set1(ray.fmax);
        for (int a = 0; a < 3; a++) {
            const typename S::T t1 = S::add(sr.ori_dir[a], S::mul(sr.rcp_dir[a], S::load(bb.mx[a])));
            const typename S::T t2 = S::add(sr.ori_dir[a], S::mul(sr.rcp_dir[a], S::load(bb.mn[a])));
            f_min = S::max(f_min, S::min(t1, t2));
            f_max = S::min(f_max, S::max(t1, t2));
        }
        const typename S::T mask = S::and_(S::load(bb.mask), S::cle(f_min, f_max));
        f_min = S::pick(mask, f_min, S::set1(-1.0f));
        return S::movemask(mask);
    }

    // simd_triangle.h:212-300 (quick_quit = false) alone: every lane's hit in (fmin, fmax] as a
    // mask, with t (FLT_MAX outside the mask), u and v per lane.
    static inline int TriLanes(const Ray& ray, const RayW<W>& sr, const TriW<W>& tri, float* tf, float* uf, float* vf) {
        using S = Simd<W>;
        using T = typename S::T;
        T mask = S::load(tri.mask);
        T p[3][3];
        for (int v = 0; v < 3; v++)
            for (int c = 0; c < 3; c++) p[v][c] = S::sub(S::load(tri.p[3 * v + c]), sr.ori[c]);
        T* p0x = &p[0][ray.kx]; T* p0y = &p[0][ray.ky]; T* p0z = &p[0][ray.kz];
        T* p1x = &p[1][ray.kx]; T* p1y = &p[1][ray.ky]; T* p1z = &p[1][ray.kz];
        T* p2x = &p[2][ray.kx]; T* p2y = &p[2][ray.ky]; T* p2z = &p[2][ray.kz];
        *p0x = S::mad(*p0y, sr.scale[0], *p0x);
        *p0z = S::mad(*p0y, sr.scale[2], *p0z);
        *p1x = S::mad(*p1y, sr.scale[0], *p1x);
        *p1z = S::mad(*p1y, sr.scale[2], *p1z);
        *p2x = S::mad(*p2y, sr.scale[0], *p2x);
        *p2z = S::mad(*p2y, sr.scale[2], *p2z);
        const T e0 = S::sub(S::mul(*p1x, *p2z), S::mul(*p1z, *p2x));
        const T e1 = S::sub(S::mul(*p2x, *p0z), S::mul(*p2z, *p0x));
        const T e2 = S::sub(S::mul(*p0x, *p1z), S::mul(*p0z, *p1x));
        const T zeros = S::set1(0.0f);
        const T c0 = S::and_(S::and_(S::cge(e0, zeros), S::cge(e1, zeros)), S::cge(e2, zeros));
        const T c1 = S::and_(S::and_(S::cle(e0, zeros), S::cle(e1, zeros)), S::cle(e2, zeros));
        mask = S::or_(S::and_(mask, c0), S::and_(mask, c1));
        if (0 == S::movemask(mask)) return 0;
        const T det = S::add(e0, S::add(e1, e2));
        mask = S::and_(mask, S::cneq(det, zeros));
        if (0 == S::movemask(mask)) return 0;
        const T rcp_det = S::rcp(det);
        *p0y = S::mul(*p0y, sr.scale[1]);
        *p1y = S::mul(*p1y, sr.scale[1]);
        *p2y = S::mul(*p2y, sr.scale[1]);
        T t = S::mul(e0, *p0y);
        t = S::mad(e1, *p1y, t);
        t = S::mad(e2, *p2y, t);
        t = S::mul(t, rcp_det);
        mask = S::and_(S::and_(mask, S::cgt(t, S::set1(ray.fmin))), S::cle(t, S::set1(ray.fmax)));
        if (0 == S::movemask(mask)) return 0;
        t = S::pick(mask, t, S::set1(FLT_MAX));
        const T u = S::mul(e1, rcp_det);
        const T v = S::mul(e2, rcp_det);
        S::store(tf, t);
        S::store(uf, u);
        S::store(vf, v);
        return S::movemask(mask);
    }

    // simd_triangle.h:212-300 (quick_quit = false) + 350-383. Returns true when a closer hit was
    // written into `hit`.
    static inline bool IntersectTri(const Ray& ray, const RayW<W>& sr, const TriW<W>& tri, Hit& hit) {
        using S = Simd<W>;
        using T = typename S::T;
        T mask = S::load(tri.mask);
        T p[3][3];
        for (int v = 0; v < 3; v++)
            for (int c = 0; c < 3; c++) p[v][c] = S::sub(S::load(tri.p[3 * v + c]), sr.ori[c]);
        T* p0x = &p[0][ray.kx]; T* p0y = &p[0][ray.ky]; T* p0z = &p[0][ray.kz];
        T* p1x = &p[1][ray.kx]; T* p1y = &p[1][ray.ky]; T* p1z = &p[1][ray.kz];
        T* p2x = &p[2][ray.kx]; T* p2y = &p[2][ray.ky]; T* p2z = &p[2][ray.kz];
        *p0x = S::mad(*p0y, sr.scale[0], *p0x);
        *p0z = S::mad(*p0y, sr.scale[2], *p0z);
        *p1x = S::mad(*p1y, sr.scale[0], *p1x);
        *p1z = S::mad(*p1y, sr.scale[2], *p1z);
        *p2x = S::mad(*p2y, sr.scale[0], *p2x);
        *p2z = S::mad(*p2y, sr.scale[2], *p2z);
        const T e0 = S::sub(S::mul(*p1x, *p2z), S::mul(*p1z, *p2x));
        const T e1 = S::sub(S::mul(*p2x, *p0z), S::mul(*p2z, *p0x));
        const T e2 = S::sub(S::mul(*p0x, *p1z), S::mul(*p0z, *p1x));
        const T zeros = S::set1(0.0f);
        const T c0 = S::and_(S::and_(S::cge(e0, zeros), S::cge(e1, zeros)), S::cge(e2, zeros));
        const T c1 = S::and_(S::and_(S::cle(e0, zeros), S::cle(e1, zeros)), S::cle(e2, zeros));
        mask = S::or_(S::and_(mask, c0), S::and_(mask, c1));
        if (0 == S::movemask(mask)) return false;
        const T det = S::add(e0, S::add(e1, e2));
        mask = S::and_(mask, S::cneq(det, zeros));
        if (0 == S::movemask(mask)) return false;
        const T rcp_det = S::rcp(det);
        *p0y = S::mul(*p0y, sr.scale[1]);
        *p1y = S::mul(*p1y, sr.scale[1]);
        *p2y = S::mul(*p2y, sr.scale[1]);
        T t = S::mul(e0, *p0y);
        t = S::mad(e1, *p1y, t);
        t = S::mad(e2, *p2y, t);
        t = S::mul(t, rcp_det);
        mask = S::and_(S::and_(mask, S::cgt(t, S::set1(ray.fmin))), S::cle(t, S::set1(ray.fmax)));
        if (0 == S::movemask(mask)) return false;
        t = S::pick(mask, t, S::set1(FLT_MAX));
        const T u = S::mul(e1, rcp_det);
        const T v = S::mul(e2, rcp_det);
        // closest selection (simd_triangle.h:359-374)
        mask = S::and_(mask, S::clt(t, S::set1(hit.t)));
        if (0 == S::movemask(mask)) return false;
        const float t0 = S::minreduce(t);
        const int resolved = S::movemask(S::ceq(t, S::set1(t0)));
        const int li = __builtin_ctz(resolved);
        alignas(32) float tf[W], uf[W], vf[W];
        S::store(tf, t);
        S::store(uf, u);
        S::store(vf, v);
        hit.t = tf[li];
        hit.u = uf[li];
        hit.v = vf[li];
        hit.prim = tri.prim[li];
        return true;
    }

    // fast_bvh.hpp:305-504 (SIMD branch). shadow = SurfaceInteraction::query_shadow.
    // Returns: closest mode -> hit found ; shadow mode -> occluded.
    // With `transp` (transparent shadows, the ENABLE_TRANSPARENT_SHADOW branches of
    // fast_bvh.hpp:352-405): per primitive id [r, g, b, has transparency]; a blocked group or quad
    // ends the walk only when its primitive has no transparency (then *opaque = true and the call
    // returns true, the reference's primitive = nullptr code); otherwise the walk goes on as a
    // closest-hit search and returns whether a hit was found.
    bool GetIntersect(Ray ray, Hit& hit, bool shadow, std::vector<std::pair<const Node<W>*, float>>& stack,
                      Stats* st, const float* transp = nullptr, bool* opaque = nullptr) const {
        hit.t = FLT_MAX;
        hit.prim = 0xffffffffu;
        hit.u = hit.v = 0.f;
        if (!valid) return false;
        if (st) st->rays++;
        ray.Prepare();
        RayW<W> sr;
        resolveRayData<W>(ray, sr);
        const float rmin = IntersectScalar(ray, sc.bbox);
        if (rmin < 0.0f) return false;
        if (stack.size() < (size_t)depth * W + W) stack.resize((size_t)depth * W + W);
        int si = 0;
        stack[si++] = std::make_pair(root.get(), rmin);
        while (si > 0) {
            const auto top = stack[--si];
            const Node<W>* node = top.first;
            const float fmin = top.second;
            if (hit.t < fmin) continue;
            if (0 == node->child_cnt) {
                if (st) st->leaves++;
                for (const auto& g : node->tris) {
                    if (st) st->tris += g.count;
                    const bool blocked = IntersectTri(ray, sr, g, hit);
                    if (shadow && blocked) {  // fast_bvh.hpp:362-373
                        if (!transp || transp[4 * (size_t)hit.prim + 3] == 0.0f) {
                            if (opaque) *opaque = true;
                            return true;
                        }
                    }
                }
                for (const uint32_t q : node->others) {
                    if (st) st->quads++;
                    float tq;
                    if (sc.quads[q].Intersect(ray, hit.t, &tq)) {
                        hit.t = tq;
                        hit.prim = sc.ntris + q;
                        hit.u = hit.v = 0.f;
                        if (shadow && (!transp || transp[4 * (size_t)hit.prim + 3] == 0.0f)) {  // :394-405
                            if (opaque) *opaque = true;
                            return true;
                        }
                    }
                }
                continue;
            }
            if (st) st->nodes++;
            typename Simd<W>::T fm;
            int m = IntersectBBox(ray, sr, node->bbox, fm);
            if (0 == m) continue;
            alignas(32) float f[W];
            Simd<W>::store(f, fm);
            const int k0 = __builtin_ctz(m);
            const float t0 = f[k0];
            m &= m - 1;
            if (0 == m) {
                stack[si++] = std::make_pair(node->children[k0].get(), t0);
            } else {
                const int k1 = __builtin_ctz(m);
                m &= m - 1;
                if (0 == m) {
                    const float t1 = f[k1];
                    if (t0 < t1) {
                        stack[si++] = std::make_pair(node->children[k1].get(), t1);
                        stack[si++] = std::make_pair(node->children[k0].get(), t0);
                    } else {
                        stack[si++] = std::make_pair(node->children[k0].get(), t0);
                        stack[si++] = std::make_pair(node->children[k1].get(), t1);
                    }
                } else {
                    for (unsigned i = 0; i < node->child_cnt; ++i) {
                        int k = -1;
                        float maxDist = -1.0f;
                        for (unsigned j = 0; j < node->child_cnt; ++j) {
                            if (f[j] > maxDist) {
                                maxDist = f[j];
                                k = (int)j;
                            }
                        }
                        if (k == -1) break;
                        f[k] = -1.0f;
                        stack[si++] = std::make_pair(node->children[k].get(), maxDist);
                    }
                }
            }
        }
        if (shadow && !transp) return false;
        return hit.prim != 0xffffffffu;
    }

    // Fbvh::GetIntersect(ray, BSSRDFIntersections&, rc, matID), fast_bvh.hpp:656-752 (SIMD branch):
    // the closest-hit walk pruned by maxt; each leaf's triangle groups through
    // intersectTriangleMulti_SIMD (simd_triangle.h:412-452): the lanes hit in the ray's window with
    // t < maxt as it was before the group, in lane order; those whose primitive has material `mat`
    // fill the four records, then replace the first record with the largest t and resolve maxt.
    // Quads are not tested (only triangles carry SSS).
    void GetIntersectMulti(Ray ray, const uint32_t* tri_mat, uint32_t mat, MultiHits& out,
                           std::vector<std::pair<const Node<W>*, float>>& stack) const {
        out.cnt = 0;
        out.maxt = FLT_MAX;
        for (auto& r : out.rec) r = Hit{FLT_MAX, 0xffffffffu, 0.f, 0.f};
        if (!valid) return;
        ray.Prepare();
        RayW<W> sr;
        resolveRayData<W>(ray, sr);
        const float rmin = IntersectScalar(ray, sc.bbox);
        if (rmin < 0.0f) return;
        if (stack.size() < (size_t)depth * W + W) stack.resize((size_t)depth * W + W);
        int si = 0;
        stack[si++] = std::make_pair(root.get(), rmin);
        while (si > 0) {
            const auto top = stack[--si];
            const Node<W>* node = top.first;
            const float fmin = top.second;
            if (out.maxt < fmin) continue;
            if (0 == node->child_cnt) {
                for (const auto& g : node->tris) {
                    alignas(32) float tf[W], uf[W], vf[W];
                    int m = TriLanes(ray, sr, g, tf, uf, vf);
                    if (!m) continue;
                    int lt = 0;
                    for (int l = 0; l < W; l++) lt |= (tf[l] < out.maxt) ? (1 << l) : 0;
                    m &= lt;
                    while (m) {
                        const int l = __builtin_ctz(m);
                        m &= m - 1;
                        const uint32_t prim = g.prim[l];
                        if (tri_mat[prim] != mat) continue;
                        const Hit h{tf[l], prim, uf[l], vf[l]};
                        if (out.cnt < 4) {
                            out.rec[out.cnt++] = h;
                        } else {
                            int picked = -1;
                            float t = 0.0f;
                            for (int i = 0; i < 4; i++)
                                if (t < out.rec[i].t) {
                                    t = out.rec[i].t;
                                    picked = i;
                                }
                            if (picked >= 0) out.rec[picked] = h;
                            // ResolveMaxDepth: std::max(a, b) = a < b ? b : a
                            float mx = out.rec[0].t < out.rec[1].t ? out.rec[1].t : out.rec[0].t;
                            mx = mx < out.rec[2].t ? out.rec[2].t : mx;
                            mx = mx < out.rec[3].t ? out.rec[3].t : mx;
                            out.maxt = mx;
                        }
                    }
                }
                continue;
            }
            typename Simd<W>::T fm;
            int m = IntersectBBox(ray, sr, node->bbox, fm);
            if (0 == m) continue;
            alignas(32) float f[W];
            Simd<W>::store(f, fm);
            const int k0 = __builtin_ctz(m);
            const float t0 = f[k0];
            m &= m - 1;
            if (0 == m) {
                stack[si++] = std::make_pair(node->children[k0].get(), t0);
            } else {
                const int k1 = __builtin_ctz(m);
                m &= m - 1;
                if (0 == m) {
                    const float t1 = f[k1];
                    if (t0 < t1) {
                        stack[si++] = std::make_pair(node->children[k1].get(), t1);
                        stack[si++] = std::make_pair(node->children[k0].get(), t0);
                    } else {
                        stack[si++] = std::make_pair(node->children[k0].get(), t0);
                        stack[si++] = std::make_pair(node->children[k1].get(), t1);
                    }
                } else {
                    for (unsigned i = 0; i < node->child_cnt; ++i) {
                        int k = -1;
                        float maxDist = -1.0f;
                        for (unsigned j = 0; j < node->child_cnt; ++j) {
                            if (f[j] > maxDist) {
                                maxDist = f[j];
                                k = (int)j;
                            }
                        }
                        if (k == -1) break;
                        f[k] = -1.0f;
                        stack[si++] = std::make_pair(node->children[k].get(), maxDist);
                    }
                }
            }
        }
    }

    // Scene::GetAttenuation (core/scene.cpp:192-208) looping Accelerator::GetAttenuation
    // (accel/accelerator.cpp:26-68). out = [r, g, b, surfaces passed]; at most `max_steps`
    // surfaces are passed (the device library's limit).
    void Attenuation(Ray ray, const float* transp, float out[4], std::vector<std::pair<const Node<W>*, float>>& stack,
                     uint32_t max_steps) const {
        float att[3] = {1.0f, 1.0f, 1.0f};
        uint32_t steps = 0;
        while (!(att[0] == 0.0f && att[1] == 0.0f && att[2] == 0.0f)) {
            Hit hit;
            bool opq = false;
            if (!GetIntersect(ray, hit, true, stack, nullptr, transp, &opq)) break;
            float a[3] = {0.0f, 0.0f, 0.0f};  // blocked: attenuation = 0
            if (!opq && transp[4 * (size_t)hit.prim + 3] != 0.0f)
                for (int c = 0; c < 3; c++) a[c] = transp[4 * (size_t)hit.prim + c];  // EvaluateTransparency
            if (a[0] == 0.0f && a[1] == 0.0f && a[2] == 0.0f) {
                out[0] = a[0]; out[1] = a[1]; out[2] = a[2]; out[3] = (float)steps;
                return;
            }
            for (int c = 0; c < 3; c++) att[c] *= a[c];
            steps++;
            if (att[0] == 0.0f && att[1] == 0.0f && att[2] == 0.0f) break;
            if (steps >= max_steps) break;
            ray.o = V3(ray.o.x + hit.t * ray.d.x, ray.o.y + hit.t * ray.d.y, ray.o.z + hit.t * ray.d.z);  // ray(t)
            ray.fmin = 0.001f;
            ray.fmax = ray.fmax - hit.t;
        }
        out[0] = att[0]; out[1] = att[1]; out[2] = att[2]; out[3] = (float)steps;
    }
};

// Scalar watertight test, shape/triangle.cpp:25-114 (the non-SIMD Obvh's per-primitive test).
// Kept as a switchable variant; differs from the SIMD test in det order, the double fallback
// and the (fMin, fMax) open window.
static inline bool ScalarTriangle(const Ray& r, const V3& op0, const V3& op1, const V3& op2, float cur_t,
                                  float* t_out, float* u_out, float* v_out) {
    V3 p0 = op0 - r.o, p1 = op1 - r.o, p2 = op2 - r.o;
    p0 = V3(p0[r.kx], p0[r.ky], p0[r.kz]);
    p1 = V3(p1[r.kx], p1[r.ky], p1[r.kz]);
    p2 = V3(p2[r.kx], p2[r.ky], p2[r.kz]);
    p0.x += r.sx * p0.y;
    p0.z += r.sz * p0.y;
    p1.x += r.sx * p1.y;
    p1.z += r.sz * p1.y;
    p2.x += r.sx * p2.y;
    p2.z += r.sz * p2.y;
    float e0 = p1.x * p2.z - p1.z * p2.x;
    float e1 = p2.x * p0.z - p2.z * p0.x;
    float e2 = p0.x * p1.z - p0.z * p1.x;
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {
        e0 = (float)((double)p1.x * (double)p2.z - (double)p1.z * (double)p2.x);
        e1 = (float)((double)p2.x * (double)p0.z - (double)p2.z * (double)p0.x);
        e2 = (float)((double)p0.x * (double)p1.z - (double)p0.z * (double)p1.x);
    }
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return false;
    const float det = e0 + e1 + e2;
    if (det == .0f) return false;
    p0.y *= r.sy;
    p1.y *= r.sy;
    p2.y *= r.sy;
    const float invDet = 1.0f / det;
    const float t = (e0 * p0.y + e1 * p1.y + e2 * p2.y) * invDet;
    if (t <= r.fmin || t >= r.fmax) return false;
    if (t > cur_t || t <= 0.0f) return false;
    *t_out = t;
    *u_out = e1 * invDet;
    *v_out = e2 * invDet;
    return true;
}

// Brute force: the same per-triangle SIMD arithmetic run over every primitive in id order,
// so the closest t is bitwise the tree's (only the order of equal-t candidates can differ).
template <int W> static bool BruteClosest(const Bvh<W>& bvh, Ray ray, Hit& hit, bool shadow) {
    hit.t = FLT_MAX;
    hit.prim = 0xffffffffu;
    hit.u = hit.v = 0.f;
    ray.Prepare();
    RayW<W> sr;
    resolveRayData<W>(ray, sr);
    const Scene& sc = bvh.sc;
    TriW<W> g;
    for (uint32_t base = 0; base < sc.ntris; base += W) {
        const uint32_t cnt = std::min<uint32_t>(W, sc.ntris - base);
        for (uint32_t l = 0; l < (uint32_t)W; l++) {
            const bool ok = l < cnt;
            g.mask[l] = ok ? mask_true() : 0.0f;
            g.prim[l] = ok ? base + l : 0xffffffffu;
            for (int k = 0; k < 3; k++) {
                const V3 p = ok ? bvh.vtx(base + l, k) : V3();
                g.p[3 * k][l] = p.x;
                g.p[3 * k + 1][l] = p.y;
                g.p[3 * k + 2][l] = p.z;
            }
        }
        g.count = cnt;
        if (Bvh<W>::IntersectTri(ray, sr, g, hit) && shadow) return true;
    }
    for (uint32_t q = 0; q < sc.quads.size(); q++) {
        float tq;
        if (sc.quads[q].Intersect(ray, hit.t, &tq)) {
            hit.t = tq;
            hit.prim = sc.ntris + q;
            hit.u = hit.v = 0.f;
            if (shadow) return true;
        }
    }
    if (shadow) return false;
    return hit.prim != 0xffffffffu;
}

}  // namespace orc

// ------------------------------------------------------------------------------------------
// C API for the tests (ctypes) and the bench's CPU baseline.
using namespace orc;

struct OracleHandle {
    int width;
    void* bvh;
};

struct RayRec {  // 32 B, same as the product's ray record
    float ox, oy, oz, tmin, dx, dy, dz, tmax;
};

static inline Ray toRay(const RayRec& r) {
    Ray ray;
    ray.o = V3(r.ox, r.oy, r.oz);
    ray.d = V3(r.dx, r.dy, r.dz);
    ray.fmin = r.tmin;
    ray.fmax = r.tmax;
    return ray;
}

template <int W>
static void fillScene(Bvh<W>* b, const float* pos, const uint32_t* idx, uint32_t ntris, const float* quads,
                      uint32_t nquads) {
    b->sc.pos = pos;
    b->sc.idx = idx;
    b->sc.ntris = ntris;
    b->sc.quads.resize(nquads);
    for (uint32_t q = 0; q < nquads; q++) {  // per quad: mat[16], inv[16], sizeX, sizeY
        const float* s = quads + 34 * q;
        std::memcpy(b->sc.quads[q].mat.m, s, 16 * sizeof(float));
        std::memcpy(b->sc.quads[q].inv.m, s + 16, 16 * sizeof(float));
        b->sc.quads[q].sizeX = s[32];
        b->sc.quads[q].sizeY = s[33];
    }
}

template <int W> static Bvh<W>* B(OracleHandle* h) { return static_cast<Bvh<W>*>(h->bvh); }

template <int W>
static void traceRange(const Bvh<W>* bvh, const RayRec* rays, uint64_t lo, uint64_t hi, Hit* hits,
                       uint8_t* occ, int mode, Stats* st) {
    std::vector<std::pair<const Node<W>*, float>> stack;
    for (uint64_t i = lo; i < hi; i++) {
        Hit h;
        if (mode == 0) {
            bvh->GetIntersect(toRay(rays[i]), h, false, stack, st);
            hits[i] = h;
        } else {
            occ[i] = bvh->GetIntersect(toRay(rays[i]), h, true, stack, st) ? 1 : 0;
        }
    }
}

template <int W>
static void dumpNode(const Node<W>* n, std::vector<uint32_t>& out) {
    out.push_back(n->child_cnt);
    out.push_back(n->pri_offset);
    out.push_back(n->pri_cnt);
    for (unsigned j = 0; j < n->child_cnt; j++) dumpNode(n->children[j].get(), out);
}

extern "C" {

// Builds the tree. quads: nquads x 34 floats (matrix, inverse, sizeX, sizeY). Arrays must
// outlive the handle.
OracleHandle* oracle_build(const float* pos, uint32_t nverts, const uint32_t* idx, uint32_t ntris,
                           const float* quads, uint32_t nquads, int width, uint32_t max_depth,
                           uint32_t max_leaf) {
    (void)nverts;
    auto* h = new OracleHandle();
    h->width = width;
    if (width == 8) {
        auto* b = new Bvh<8>();
        b->maxDepth = max_depth;
        b->maxLeaf = max_leaf;
        fillScene<8>(b, pos, idx, ntris, quads, nquads);
        b->Build();
        h->bvh = b;
    } else {
        auto* b = new Bvh<4>();
        b->maxDepth = max_depth;
        b->maxLeaf = max_leaf;
        fillScene<4>(b, pos, idx, ntris, quads, nquads);
        b->Build();
        h->bvh = b;
    }
    return h;
}

void oracle_destroy(OracleHandle* h) {
    if (!h) return;
    if (h->width == 8) delete B<8>(h);
    else delete B<4>(h);
    delete h;
}

// info: [depth, nodeCount, leafCount, nprims, valid]
void oracle_info(OracleHandle* h, uint64_t* info) {
    if (h->width == 8) {
        auto* b = B<8>(h);
        info[0] = b->depth; info[1] = b->nodeCount; info[2] = b->leafCount; info[3] = b->prims.size(); info[4] = b->valid;
    } else {
        auto* b = B<4>(h);
        info[0] = b->depth; info[1] = b->nodeCount; info[2] = b->leafCount; info[3] = b->prims.size(); info[4] = b->valid;
    }
}

void oracle_scene_bbox(OracleHandle* h, float* out6) {
    const BBox& b = h->width == 8 ? B<8>(h)->sc.bbox : B<4>(h)->sc.bbox;
    for (int i = 0; i < 3; i++) {
        out6[i] = b.mn[i];
        out6[3 + i] = b.mx[i];
    }
}

// Primitive order after the build (m_bvhpri), length = nprims.
void oracle_prim_order(OracleHandle* h, uint32_t* out) {
    if (h->width == 8) {
        auto* b = B<8>(h);
        for (size_t i = 0; i < b->prims.size(); i++) out[i] = b->prims[i].id;
    } else {
        auto* b = B<4>(h);
        for (size_t i = 0; i < b->prims.size(); i++) out[i] = b->prims[i].id;
    }
}

// Pre-order dump of (child_cnt, pri_offset, pri_cnt) triples. Returns the number of uint32
// written (call with out=nullptr to size).
uint64_t oracle_tree_dump(OracleHandle* h, uint32_t* out, uint64_t cap) {
    std::vector<uint32_t> v;
    if (h->width == 8) {
        if (B<8>(h)->root) dumpNode<8>(B<8>(h)->root.get(), v);
    } else {
        if (B<4>(h)->root) dumpNode<4>(B<4>(h)->root.get(), v);
    }
    if (out) std::memcpy(out, v.data(), std::min<uint64_t>(cap, v.size()) * 4);
    return v.size();
}

// mode 0: closest (hits: 16 B records); mode 1: any-hit (occ: u8). stats[5] accumulates
// rays, inner nodes, leaves, triangles, quads. nthreads <= 1 runs on the calling thread.
// Returns wall seconds.
double oracle_trace(OracleHandle* h, const void* rays_, uint64_t n, void* hits_, uint8_t* occ, int mode,
                    int nthreads, uint64_t* stats) {
    const RayRec* rays = static_cast<const RayRec*>(rays_);
    Hit* hits = static_cast<Hit*>(hits_);
    const auto t0 = std::chrono::steady_clock::now();
    if (nthreads <= 1) {
        Stats st;
        if (h->width == 8) traceRange<8>(B<8>(h), rays, 0, n, hits, occ, mode, &st);
        else traceRange<4>(B<4>(h), rays, 0, n, hits, occ, mode, &st);
        if (stats) {
            stats[0] += st.rays; stats[1] += st.nodes; stats[2] += st.leaves; stats[3] += st.tris; stats[4] += st.quads;
        }
    } else {
        // worker pool pulling 4096-ray chunks
        std::atomic<uint64_t> next{0};
        const uint64_t chunk = 4096;
        std::vector<Stats> sts(nthreads);
        std::vector<std::thread> th;
        for (int w = 0; w < nthreads; w++) {
            th.emplace_back([&, w]() {
                Stats* st = stats ? &sts[w] : nullptr;
                while (true) {
                    const uint64_t lo = next.fetch_add(chunk);
                    if (lo >= n) break;
                    const uint64_t hi = std::min(n, lo + chunk);
                    if (h->width == 8) traceRange<8>(B<8>(h), rays, lo, hi, hits, occ, mode, st);
                    else traceRange<4>(B<4>(h), rays, lo, hi, hits, occ, mode, st);
                }
            });
        }
        for (auto& t : th) t.join();
        if (stats)
            for (auto& st : sts) {
                stats[0] += st.rays; stats[1] += st.nodes; stats[2] += st.leaves; stats[3] += st.tris; stats[4] += st.quads;
            }
    }
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// One ray on the calling thread with a per-thread stack (the path tracer's per-vertex queries,
// like a render thread calling Scene::GetIntersect). mode 0 closest (hit16), 1 any (returns 1
// when occluded).
int oracle_ray1(OracleHandle* h, const void* ray_, void* hit_, int mode) {
    const RayRec* r = static_cast<const RayRec*>(ray_);
    Hit hh;
    bool res;
    if (h->width == 8) {
        thread_local std::vector<std::pair<const Node<8>*, float>> st;
        res = B<8>(h)->GetIntersect(toRay(*r), hh, mode == 1, st, nullptr);
    } else {
        thread_local std::vector<std::pair<const Node<4>*, float>> st;
        res = B<4>(h)->GetIntersect(toRay(*r), hh, mode == 1, st, nullptr);
    }
    if (mode == 0) *static_cast<Hit*>(hit_) = hh;
    return res ? 1 : 0;
}

// Transparent-shadow attenuation of one ray on the calling thread (the path tracer's shadow
// queries in a scene with transparent materials): out = (r, g, b, surfaces passed).
void oracle_atten1(OracleHandle* h, const void* ray_, const float* transp, float* out, uint32_t max_steps) {
    const RayRec* r = static_cast<const RayRec*>(ray_);
    if (h->width == 8) {
        thread_local std::vector<std::pair<const Node<8>*, float>> st;
        B<8>(h)->Attenuation(toRay(*r), transp, out, st, max_steps);
    } else {
        thread_local std::vector<std::pair<const Node<4>*, float>> st;
        B<4>(h)->Attenuation(toRay(*r), transp, out, st, max_steps);
    }
}

// Transparent-shadow attenuation of n rays: transp = 4 floats per primitive id; out = 4 floats per
// ray (r, g, b, surfaces passed). Walks stop after max_steps surfaces.
void oracle_attenuation(OracleHandle* h, const void* rays_, uint64_t n, const float* transp, float* out,
                        uint32_t max_steps, int nthreads) {
    const RayRec* rays = static_cast<const RayRec*>(rays_);
    std::atomic<uint64_t> next{0};
    auto work = [&]() {
        std::vector<std::pair<const Node<8>*, float>> s8;
        std::vector<std::pair<const Node<4>*, float>> s4;
        while (true) {
            const uint64_t lo = next.fetch_add(1024);
            if (lo >= n) break;
            const uint64_t hi = std::min(n, lo + 1024);
            for (uint64_t i = lo; i < hi; i++) {
                if (h->width == 8) B<8>(h)->Attenuation(toRay(rays[i]), transp, out + 4 * i, s8, max_steps);
                else B<4>(h)->Attenuation(toRay(rays[i]), transp, out + 4 * i, s4, max_steps);
            }
        }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < nthreads; w++) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

// BSSRDF multi-hit query of n rays: ray_mat[i] = the material ray i looks for, tri_mat = material
// per triangle id; hits = 4 records per ray (t, prim, u, v; unused records FLT_MAX / 0xffffffff),
// cnt = records filled.
void oracle_multi(OracleHandle* h, const void* rays_, uint64_t n, const uint32_t* ray_mat, const uint32_t* tri_mat,
                  void* hits_, uint32_t* cnt, int nthreads) {
    const RayRec* rays = static_cast<const RayRec*>(rays_);
    Hit* hits = static_cast<Hit*>(hits_);
    std::atomic<uint64_t> next{0};
    auto work = [&]() {
        std::vector<std::pair<const Node<8>*, float>> s8;
        std::vector<std::pair<const Node<4>*, float>> s4;
        MultiHits mh;
        while (true) {
            const uint64_t lo = next.fetch_add(1024);
            if (lo >= n) break;
            const uint64_t hi = std::min(n, lo + 1024);
            for (uint64_t i = lo; i < hi; i++) {
                if (h->width == 8) B<8>(h)->GetIntersectMulti(toRay(rays[i]), tri_mat, ray_mat[i], mh, s8);
                else B<4>(h)->GetIntersectMulti(toRay(rays[i]), tri_mat, ray_mat[i], mh, s4);
                for (int k = 0; k < 4; k++) hits[4 * i + k] = mh.rec[k];
                cnt[i] = mh.cnt;
            }
        }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < nthreads; w++) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

// Brute force over all primitives (same per-triangle arithmetic). mode as above.
void oracle_brute(OracleHandle* h, const void* rays_, uint64_t n, void* hits_, uint8_t* occ, int mode) {
    const RayRec* rays = static_cast<const RayRec*>(rays_);
    Hit* hits = static_cast<Hit*>(hits_);
    for (uint64_t i = 0; i < n; i++) {
        Hit hh;
        bool r;
        if (h->width == 8) r = BruteClosest<8>(*B<8>(h), toRay(rays[i]), hh, mode == 1);
        else r = BruteClosest<4>(*B<4>(h), toRay(rays[i]), hh, mode == 1);
        if (mode == 0) hits[i] = hh;
        else occ[i] = r ? 1 : 0;
    }
}

// Ray::Prepare + resolveRayData outputs for one ray: [kx,ky,kz] ints, then
// [sx,sy,sz, rcp_x,rcp_y,rcp_z, oridir_x,oridir_y,oridir_z] floats.
void oracle_ray_setup(const void* ray_, int32_t* k3, float* f9) {
    Ray r = toRay(*static_cast<const RayRec*>(ray_));
    r.Prepare();
    RayW<8> s;
    resolveRayData<8>(r, s);
    k3[0] = r.kx; k3[1] = r.ky; k3[2] = r.kz;
    f9[0] = r.sx; f9[1] = r.sy; f9[2] = r.sz;
    for (int i = 0; i < 3; i++) {
        alignas(32) float a[8], b[8];
        _mm256_store_ps(a, s.rcp_dir[i]);
        _mm256_store_ps(b, s.ori_dir[i]);
        f9[3 + i] = a[0];
        f9[6 + i] = b[0];
    }
}

// One triangle against one ray with the SIMD arithmetic (lane 0 of a group). Returns 1 and
// writes t,u,v on a hit inside (tmin, tmax].
int oracle_tri_simd(const void* ray_, const float* tri9, float* tuv) {
    Ray r = toRay(*static_cast<const RayRec*>(ray_));
    r.Prepare();
    RayW<8> s;
    resolveRayData<8>(r, s);
    TriW<8> g;
    for (int l = 0; l < 8; l++) {
        g.mask[l] = l == 0 ? mask_true() : 0.0f;
        for (int c = 0; c < 9; c++) g.p[c][l] = l == 0 ? tri9[c] : 0.0f;
        g.prim[l] = l == 0 ? 0 : 0xffffffffu;
    }
    g.count = 1;
    Hit h{FLT_MAX, 0xffffffffu, 0.f, 0.f};
    if (!Bvh<8>::IntersectTri(r, s, g, h)) return 0;
    tuv[0] = h.t; tuv[1] = h.u; tuv[2] = h.v;
    return 1;
}

// Scalar Triangle::GetIntersect for one triangle. Returns 1 and writes t,u,v on a hit.
int oracle_tri_scalar(const void* ray_, const float* tri9, float* tuv) {
    Ray r = toRay(*static_cast<const RayRec*>(ray_));
    r.Prepare();
    float t, u, v;
    if (!ScalarTriangle(r, V3(tri9[0], tri9[1], tri9[2]), V3(tri9[3], tri9[4], tri9[5]),
                        V3(tri9[6], tri9[7], tri9[8]), FLT_MAX, &t, &u, &v))
        return 0;
    tuv[0] = t; tuv[1] = u; tuv[2] = v;
    return 1;
}

// 8 boxes (mn[3][8], mx[3][8], valid[8] as 0/1) against one ray. Returns the movemask and
// writes the 8 entry distances.
int oracle_bbox8(const void* ray_, const float* boxes48, const int32_t* valid8, float* fmin8) {
    Ray r = toRay(*static_cast<const RayRec*>(ray_));
    r.Prepare();
    RayW<8> s;
    resolveRayData<8>(r, s);
    BBoxW<8> bb;
    for (int a = 0; a < 3; a++)
        for (int l = 0; l < 8; l++) {
            bb.mn[a][l] = boxes48[a * 8 + l];
            bb.mx[a][l] = boxes48[24 + a * 8 + l];
        }
    for (int l = 0; l < 8; l++) bb.mask[l] = valid8[l] ? mask_true() : 0.0f;
    __m256 fm;
    const int m = Bvh<8>::IntersectBBox(r, s, bb, fm);
    _mm256_storeu_ps(fmin8, fm);
    return m;
}

float oracle_root_slab(const void* ray_, const float* box6) {
    Ray r = toRay(*static_cast<const RayRec*>(ray_));
    BBox b;
    b.mn = V3(box6[0], box6[1], box6[2]);
    b.mx = V3(box6[3], box6[4], box6[5]);
    return IntersectScalar(r, b);
}

// Quad test (quad.cpp:66-94) with limit FLT_MAX. q34 = mat, inv, sizeX, sizeY.
int oracle_quad(const void* ray_, const float* q34, float* t) {
    Ray r = toRay(*static_cast<const RayRec*>(ray_));
    Quad q;
    std::memcpy(q.mat.m, q34, 64);
    std::memcpy(q.inv.m, q34 + 16, 64);
    q.sizeX = q34[32];
    q.sizeY = q34[33];
    return q.Intersect(r, FLT_MAX, t) ? 1 : 0;
}

// Known-answer helpers for the SIMD wrapper semantics (unittests/simd.hpp).
void oracle_simd_rcp8(const float* in8, float* out8) {
    _mm256_storeu_ps(out8, Simd<8>::rcp(_mm256_loadu_ps(in8)));
}
float oracle_simd_minreduce8(const float* in8) { return Simd<8>::minreduce(_mm256_loadu_ps(in8)); }
int oracle_simd_movemask8(const float* in8) { return _mm256_movemask_ps(_mm256_loadu_ps(in8)); }

}  // extern "C"
