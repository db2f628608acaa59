// BSSRDF multi-hit query: up to four triangle hits of one material along a ray, with the
// reference's record update (Fbvh::GetIntersect with BSSRDFIntersections, fast_bvh.hpp:656-752;
// intersectTriangleMulti_SIMD, simd_triangle.h:412-452; BSSRDFIntersections, bssrdf.h:32-56).
//
// One ray per lane with its own stack (private memory): subsurface probe rays are a small share of
// a frame's rays and end after a few leaves, so this kernel keeps the closest-hit kernel's slab
// and triangle arithmetic and push order, without its wave-level refill and leaf batching.
#include <hip/hip_runtime.h>

#include "trace_launch.h"
#include "traverse.h"

namespace sortgpu {
namespace {

constexpr int kMultiBlock = 256;

template <int W, int STACK>
__global__ __launch_bounds__(kMultiBlock) void k_multi(MultiArgs a) {
    const uint64_t slot_k = (uint64_t)blockIdx.x * kMultiBlock + threadIdx.x;
    if (slot_k >= a.n) return;
    // the k-th probe of the traversal order (a permutation of 0..n-1), results at its own index
    const uint64_t i = a.order ? a.order[slot_k] : slot_k;
    const NodeRec<W>* __restrict__ nodes = static_cast<const NodeRec<W>*>(a.tree.nodes);
    const PrimRec* __restrict__ prims = a.tree.prims;
    const float4* rp = reinterpret_cast<const float4*>(a.rays) + 2 * i;
    const float4 q0 = rp[0], q1 = rp[1];
    RayK r;
    r.ox = q0.x; r.oy = q0.y; r.oz = q0.z; r.tmin = q0.w;
    r.dx = q1.x; r.dy = q1.y; r.dz = q1.z; r.tmax = q1.w;
    ray_setup(r);
    r.ieee = ieee_slabs_ok(r, a.tree.bbox);
    const uint32_t mat = a.ray_material[i];

    // the four records (t, prim, u, v), filled in order, then replaced at the first largest t
    float rt0 = FLT_MAX, rt1 = FLT_MAX, rt2 = FLT_MAX, rt3 = FLT_MAX;
    uint32_t rp0 = 0xffffffffu, rp1 = 0xffffffffu, rp2 = 0xffffffffu, rp3 = 0xffffffffu;
    float ru0 = 0.f, ru1 = 0.f, ru2 = 0.f, ru3 = 0.f, rv0 = 0.f, rv1 = 0.f, rv2 = 0.f, rv3 = 0.f;
    uint32_t cnt = 0;
    float maxt = FLT_MAX;  // resolved only when a record is replaced (bssrdf.h:50-55)

    uint2 st[STACK];
    int sp = 0;
    bool overflow = false;
    if (a.tree.valid) {
        const float rmin = root_slab(r, a.tree.bbox);
        if (rmin >= 0.0f) st[sp++] = make_uint2(a.tree.root, __float_as_uint(rmin));
    }
    while (sp > 0) {
        const uint2 e = st[--sp];
        const uint32_t ref = e.x;
        if (maxt < __uint_as_float(e.y)) continue;
        if (ref & REF_LEAF) {
            // triangle groups of W (PR_GEND ends one): every hit of a group is judged against maxt
            // as it was before the group; quads are not tested. The leaf's next record loads while
            // the current one is tested.
            uint32_t li = ref & ~REF_LEAF;
            bool gstart = true;
            float gmaxt = maxt;
            auto tri = [&](const Tri& x, uint32_t id) {
                if (gstart) {
                    gmaxt = maxt;
                    gstart = false;
                }
                const uint32_t prim = id & PR_ID;
                float t, u, v;
                if (tri_test(r, x, t, u, v) && t < gmaxt && a.tri_material[prim] == mat) {
                    int slot;
                    const bool full = cnt >= 4u;
                    if (!full) {
                        slot = (int)cnt++;
                    } else {
                        // the first record with the largest t (strictly above 0)
                        slot = -1;
                        float mt = 0.0f;
                        if (mt < rt0) { mt = rt0; slot = 0; }
                        if (mt < rt1) { mt = rt1; slot = 1; }
                        if (mt < rt2) { mt = rt2; slot = 2; }
                        if (mt < rt3) { mt = rt3; slot = 3; }
                    }
                    if (slot == 0) { rt0 = t; rp0 = prim; ru0 = u; rv0 = v; }
                    if (slot == 1) { rt1 = t; rp1 = prim; ru1 = u; rv1 = v; }
                    if (slot == 2) { rt2 = t; rp2 = prim; ru2 = u; rv2 = v; }
                    if (slot == 3) { rt3 = t; rp3 = prim; ru3 = u; rv3 = v; }
                    if (full) maxt = smax(smax(smax(rt0, rt1), rt2), rt3);
                }
                if (id & PR_GEND) gstart = true;
            };
            PrimRec pr = load_prim(prims, li);
            while (true) {
                const bool last = (pr.id_b & PR_FLAG) != 0;
                PrimRec nxt = pr;
                if (!last) nxt = load_prim(prims, li + 1);
                if (!(pr.id_a & PR_FLAG)) {
                    tri(pr.a, pr.id_a);
                    if ((pr.id_b & PR_ID) != PR_NONE) tri(pr.b, pr.id_b);
                }
                if (last) break;
                pr = nxt;
                ++li;
            }
            continue;
        }
        float f[W];
        uint32_t ch[W];
        const uint32_t m = node_test<W>(r, nodes, ref, f, ch);
        uint32_t elig = m;
        if (__popc(m) >= 3) {
            // repeated argmax over f > -1 (fast_bvh.hpp:725-740)
#pragma unroll
            for (int j = 0; j < W; j++)
                if (!(f[j] > -1.0f)) elig &= ~(1u << j);
        }
        // farther first, ties lower lane first (the hit children's distances are not NaN)
        uint32_t rank[W];
#pragma unroll
        for (int j = 0; j < W; j++) rank[j] = 0;
#pragma unroll
        for (int j = 1; j < W; j++) {
#pragma unroll
            for (int k = 0; k < j; k++) {
                const bool kfirst = f[k] >= f[j];
                rank[j] += kfirst ? ((elig >> k) & 1u) : 0u;
                rank[k] += kfirst ? 0u : ((elig >> j) & 1u);
            }
        }
        const int n = __popc(elig);
        if (sp + n > STACK) {  // cannot happen for a tree the host accepted (max_stack <= STACK)
            overflow = true;
            break;
        }
#pragma unroll
        for (int j = 0; j < W; j++)
            if (elig & (1u << j)) st[sp + (int)rank[j]] = make_uint2(ch[j], __float_as_uint(f[j]));
        sp += n;
    }
    if (overflow) atomicOr(a.error, kFlagStackOverflow);
    float4* out = reinterpret_cast<float4*>(a.hits) + 4 * i;
    out[0] = make_float4(rt0, __uint_as_float(rp0), ru0, rv0);
    out[1] = make_float4(rt1, __uint_as_float(rp1), ru1, rv1);
    out[2] = make_float4(rt2, __uint_as_float(rp2), ru2, rv2);
    out[3] = make_float4(rt3, __uint_as_float(rp3), ru3, rv3);
    a.counts[i] = cnt;
}

template <int W, int STACK>
hipError_t launch_multi_w(const MultiArgs& a, hipStream_t s) {
    const uint64_t blocks = (a.n + kMultiBlock - 1) / kMultiBlock;
    hipLaunchKernelGGL((k_multi<W, STACK>), dim3((uint32_t)blocks), dim3(kMultiBlock), 0, s, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_multi(const MultiArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    const uint32_t need = a.tree.max_stack;
    if (a.tree.width == 8) {
        if (need <= 64) return launch_multi_w<8, 64>(a, s);
        if (need <= 256) return launch_multi_w<8, 256>(a, s);
    } else {
        if (need <= 64) return launch_multi_w<4, 64>(a, s);
        if (need <= 256) return launch_multi_w<4, 256>(a, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace sortgpu
