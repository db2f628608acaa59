// Closest-hit and any-hit traversal kernels over flattened Fast-BVHs (8- or 4-wide).
//
// Execution model: persistent waves, one ray per lane. A lane whose stack empties writes its
// result and, together with the other idle lanes of its wave, takes the next rays from a
// device work counter (one atomic per refill, taken when at least `refill` lanes are idle or
// the whole wave is), so long rays do not hold a wave's other 63 lanes idle.
//
// Multi-scene form (MULTI): every ray names its scene (ray_scene[i]); a lane loads that
// scene's tree view (root, box, record pointers into the concatenated pool) when it takes the
// ray, so rays of any number of scenes share one launch and one work queue.
//
// Traversal stack: 8-byte entries (child reference, entry distance), so a pop is one LDS read.
// The top LDS entries of every lane live in LDS ([entry][thread], so a wave's accesses to one
// entry hit 64 consecutive entries); deeper entries spill to a private per-lane array in
// scratch. The host picks the spill depth from the tree's exact worst case (W-1)*(depth-1)+1, so
// no push can go past it; the kernel still bounds-checks every push and raises an error flag
// instead of writing.
//
// The walk is SORT's Fbvh::GetIntersect (accel/fast_bvh.hpp:305-504): pop, prune when the
// current closest t is below the entry distance, test the leaf's primitives in order, or test
// the W child boxes and push the hit children farthest first (ties: lower lane pushed first).
// Per wave the loop alternates a node phase (walking lanes step inner nodes together) and a leaf
// phase (lanes holding a leaf test one primitive record each per step); a lane loads its leaf's
// next record while it tests the current one. Any-hit pushes eligible children in lane order:
// with no closest-hit pruning before the first hit, occlusion does not depend on visit order.
// Slab tests of rays with finite components against trees with bounded boxes use IEEE
// min/max (equal to the x86 selects when no operand is NaN, which such rays cannot produce);
// other rays take the select form (traverse.h, ieee_slabs_ok). An 8-wide node test loads the
// second half of each plane only when the node has more than four children (traverse.h,
// node_test): the 16-byte loads each lane issues are what bounds the walk.
#include <hip/hip_runtime.h>

#include "traverse.h"
#include "trace_launch.h"

#pragma clang fp contract(off)

namespace sortgpu {

constexpr int kBlock = 256;
constexpr int kDefaultLds = 16;
constexpr uint32_t kDefaultLeafBatch = 32;

// One triangle of a leaf record, in record order. Closest hit keeps the first triangle with the
// smallest t in (tmin, tmax] (strict t < bt); any hit ends the ray at a hit. Transparent-shadow
// walk (TRANSP): a closest-hit walk that stops, blocked, when a group of W triangles (PR_GEND
// ends one) makes the closest hit and that primitive's material has no transparency
// (fast_bvh.hpp:352-373, the ENABLE_TRANSPARENT_SHADOW branch: only the group's nearest hit is
// checked). `gup`: the current group has made the closest hit. Returns true when the walk stops.
template <bool ANY, bool STATS, bool TRANSP>
__device__ __forceinline__ bool tri_step(const RayK& r, const Tri& tri, uint32_t id, const float* __restrict__ transp,
                                         float& bt, uint32_t& bp, float& bu, float& bv, bool& gup,
                                         uint32_t& n_tris) {
    if (STATS) n_tris++;
    float t, u, v;
    bool hit = false;
    if (tri_test(r, tri, t, u, v) && t < bt) {
        bt = t;
        bp = id & PR_ID;
        bu = u;
        bv = v;
        hit = true;
    }
    if (TRANSP) {
        gup |= hit;
        if (id & PR_GEND) {
            const bool blocked = gup && (!transp || transp[4 * (size_t)bp + 3] == 0.0f);
            gup = false;
            return blocked;
        }
        return false;
    }
    return ANY && hit;
}

// One leaf record of a lane's leaf: `pr` is tested while the leaf's next record (if any) is
// loaded into `nxt`: its triangles A then B, or its quad (a quad at the current closest t
// replaces it; in the transparent-shadow walk a quad that makes the closest hit blocks when its
// material has no transparency, fast_bvh.hpp:394-405).
template <bool ANY, bool STATS, bool TRANSP>
__device__ __forceinline__ void leaf_step(const RayK& r, const PrimRec* __restrict__ prims,
                                          const QuadRec* __restrict__ quads, const float* __restrict__ transp,
                                          const PrimRec& pr, PrimRec& nxt, uint32_t& li, int& sp, float& bt,
                                          uint32_t& bp, float& bu, float& bv, bool& occluded, bool& gup,
                                          uint32_t& n_tris, uint32_t& n_quads) {
    if (li == 0xffffffffu) return;
    const bool last = (pr.id_b & PR_FLAG) != 0;
    if (!last) nxt = load_prim(prims, li + 1);
    bool stop = false;
    if (pr.id_a & PR_FLAG) {
        if (STATS) n_quads++;
        float t;
        if (quad_test(r, quads[__float_as_uint(pr.a.p0x)], bt, t)) {
            bt = t;
            bp = pr.id_a & PR_ID;
            bu = bv = 0.f;
            stop = ANY || (TRANSP && (!transp || transp[4 * (size_t)bp + 3] == 0.0f));
        }
    } else {
        stop = tri_step<ANY, STATS, TRANSP>(r, pr.a, pr.id_a, transp, bt, bp, bu, bv, gup, n_tris);
        if (!stop && (pr.id_b & PR_ID) != PR_NONE)
            stop = tri_step<ANY, STATS, TRANSP>(r, pr.b, pr.id_b, transp, bt, bp, bu, bv, gup, n_tris);
    }
    if (stop) {
        occluded = true;
        li = 0xffffffffu;
        sp = 0;
    } else {
        li = last ? 0xffffffffu : li + 1;
    }
}

template <int W, bool ANY, bool STATS, int SPILL, bool MULTI, int LDS, bool TRANSP = false>
// At least 3 waves per SIMD (at most 168 VGPRs): the occupancy every W = 8 form is tuned for.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_trace(TraceArgs a) {
    // stack entry: (child reference, entry distance bits)
    __shared__ uint2 s_st[LDS * kBlock];
    constexpr int CAP = LDS + SPILL;
    uint2 sp_st[SPILL];

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t n_rays = a.n_dev ? *a.n_dev : a.n;
    const uint32_t refill = a.refill ? a.refill : 16u;
    const uint32_t leaf_batch = a.leaf_batch ? a.leaf_batch : kDefaultLeafBatch;
    const uint32_t leaf_min = a.leaf_min ? a.leaf_min : 1u;
    const NodeRec<W>* __restrict__ nodes = static_cast<const NodeRec<W>*>(a.tree.nodes);
    const PrimRec* __restrict__ prims = a.tree.prims;
    const QuadRec* __restrict__ quads = a.tree.quads;
    const float* __restrict__ transp = a.transp;
    const uint32_t ostride = a.out_stride ? a.out_stride : 1u;  // attenuation output stride (float4s)
    uint32_t scene = 0;  // multi-scene transparent walk: the lane's tree index (restarts reread its view)

    RayK r;
    uint32_t li = 0xffffffffu;  // next primitive record of this lane's leaf
    PrimRec cur, alt;           // that record, loaded ahead; the leaf phase alternates the two
    uint32_t ray_id = 0xffffffffu;
    bool more = true;
    int sp = 0;
    float bt = FLT_MAX, bu = 0.f, bv = 0.f;
    uint32_t bp = 0xffffffffu;
    uint32_t n_nodes = 0, n_leaves = 0, n_tris = 0, n_quads = 0, n_wide = 0;
    bool overflow = false;

    auto push = [&](int e, uint32_t ref, float t) {
        const uint2 v = make_uint2(ref, __float_as_uint(t));
        if (e < LDS) {
            s_st[e * kBlock + tid] = v;
        } else if (e < CAP) {
            sp_st[e - LDS] = v;
        } else {
            overflow = true;
        }
    };

    bool occluded = false;
    // transparent-shadow walk: the current group made the closest hit; attenuation so far; surfaces passed
    bool gup = false;
    float att_r = 1.0f, att_g = 1.0f, att_b = 1.0f;
    uint32_t steps = 0;
    // Whether the leaf phase takes another step. It may stop while leaves are unfinished only
    // when fewer than `leaf_min` lanes hold one and the next node phase can take a step (some lane
    // can walk and fewer than `leaf_batch` lanes hold a leaf), so every round advances some lane
    // whatever the knobs.
    auto leaf_go = [&]() -> bool {
        const uint64_t holding = __ballot(li != 0xffffffffu);
        if (holding == 0) return false;
        const uint32_t cnt = (uint32_t)__popcll(holding);
        return !(cnt < leaf_min && cnt < leaf_batch && __ballot(li == 0xffffffffu && sp > 0) != 0);
    };

    while (true) {
        const bool idle = sp == 0 && li == 0xffffffffu;
        const uint64_t want = __ballot(idle && more);
        const uint64_t busy = __ballot(!idle);
        if (want != 0 && (busy == 0 || (uint32_t)__popcll(want) >= refill)) {
            const uint32_t cnt = (uint32_t)__popcll(want);
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)want) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(a.counter, cnt);
            base = __shfl(base, leader);
            if (idle && more) {
                const uint32_t k = base + (uint32_t)__popcll(want & ((1ull << lane) - 1ull));
                if (k < n_rays) {
                    // the k-th ray of the traversal order (order has n entries, k < n)
                    const uint32_t my = a.order ? a.order[k] : k;
                    ray_id = my;
                    const float4* rp = reinterpret_cast<const float4*>(a.rays) + 2 * (size_t)my;
                    const float4 q0 = rp[0], q1 = rp[1];
                    r.ox = q0.x; r.oy = q0.y; r.oz = q0.z; r.tmin = q0.w;
                    r.dx = q1.x; r.dy = q1.y; r.dz = q1.z; r.tmax = q1.w;
                    ray_setup(r);
                    bt = FLT_MAX;
                    bp = 0xffffffffu;
                    bu = bv = 0.f;
                    if (TRANSP) {
                        att_r = att_g = att_b = 1.0f;
                        steps = 0;
                        gup = false;
                    }
                    uint32_t root, valid;
                    float rmin = -1.0f;
                    if (MULTI) {
                        const uint32_t sid = a.ray_scene[my];
                        if (sid < a.n_trees) {
                            const TreeView& tv = a.trees[sid];
                            nodes = static_cast<const NodeRec<W>*>(tv.nodes);
                            prims = tv.prims;
                            quads = tv.quads;
                            if (TRANSP) {
                                scene = sid;
                                transp = a.transps[sid];
                            }
                            root = tv.root;
                            valid = tv.valid;
                            if (valid) rmin = root_slab(r, tv.bbox);
                            r.ieee = ieee_slabs_ok(r, tv.bbox);
                        } else {
                            root = 0;
                            valid = 0;
                            atomicOr(a.error, kFlagBadScene);
                        }
                    } else {
                        root = a.tree.root;
                        valid = a.tree.valid;
                        if (valid) rmin = root_slab(r, a.tree.bbox);
                        r.ieee = ieee_slabs_ok(r, a.tree.bbox);
                    }
                    if (rmin >= 0.0f) {
                        push(0, root, rmin);
                        sp = 1;
                    } else {
                        if (TRANSP) reinterpret_cast<float4*>(a.out)[(size_t)my * ostride] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
                        else if (ANY) {
                            if (!a.occ_slot) static_cast<uint8_t*>(a.out)[my] = 0;
                        }
                        else reinterpret_cast<float4*>(a.out)[my] =
                                 make_float4(FLT_MAX, __uint_as_float(0xffffffffu), 0.f, 0.f);
                        ray_id = 0xffffffffu;
                    }
                } else {
                    more = false;
                }
            }
        }
        if (__ballot(sp > 0 || li != 0xffffffffu) == 0) {
            if (__ballot(more) == 0) break;
            continue;
        }
        // Node phase: lanes without a leaf walk inner nodes (pop, prune, test the children, push)
        // until they pop a leaf or their stack empties; it ends when no lane can walk or
        // `leaf_batch` lanes hold a leaf, so node steps of many lanes run together.
        while (true) {
            const bool walk = li == 0xffffffffu && sp > 0;
            if (__ballot(walk) == 0 || (uint32_t)__popcll(__ballot(li != 0xffffffffu)) >= leaf_batch) break;
            if (walk) {
                --sp;
                // the LDS read is unconditional (clamped index) so that it stays an LDS access
                uint2 e = s_st[(sp < LDS ? sp : LDS - 1) * kBlock + tid];
                if (sp >= LDS) e = sp_st[sp - LDS < SPILL ? sp - LDS : 0];
                const uint32_t ref = e.x;
                const float tm = __uint_as_float(e.y);
                if (!(bt < tm)) {
                    if (ref & REF_LEAF) {
                        if (STATS) n_leaves++;
                        li = ref & ~REF_LEAF;
                        cur = load_prim(prims, li);
                    } else {
                        if (STATS) n_nodes++;
                        float f[W];
                        uint32_t ch[W];
                        const uint32_t m = node_test<W>(r, nodes, ref, f, ch);
                        if (STATS && W == 8 && ch[4 % W] != REF_EMPTY) n_wide++;
                        const int nh = __popc(m);
                        uint32_t elig = m;
                        if (nh >= 3) {
                            // the >=3 branch pushes by repeated argmax over f > -1 (fast_bvh.hpp:438-453)
#pragma unroll
                            for (int j = 0; j < W; j++)
                                if (!(f[j] > -1.0f)) elig &= ~(1u << j);
                        }
                        // Push order: farther first, ties lower lane first. An eligible child hit its
                        // box (fmn <= fmx), so its distance is not NaN and the order is total: one
                        // comparison per pair decides which of the two goes first.
                        uint32_t rank[W];
#pragma unroll
                        for (int j = 0; j < W; j++) rank[j] = 0;
                        if (!(ANY && !STATS)) {
#pragma unroll
                            for (int j = 1; j < W; j++) {
#pragma unroll
                                for (int k = 0; k < j; k++) {
                                    const bool kfirst = f[k] >= f[j];
                                    rank[j] += kfirst ? ((elig >> k) & 1u) : 0u;
                                    rank[k] += kfirst ? 0u : ((elig >> j) & 1u);
                                }
                            }
                        }
#pragma unroll
                        for (int j = 0; j < W; j++) {
                            if (elig & (1u << j)) {
                                const uint32_t rk = (ANY && !STATS) ? (uint32_t)__popc(elig & ((1u << j) - 1u)) : rank[j];
                                push(sp + (int)rk, ch[j], f[j]);
                            }
                        }
                        sp += __popc(elig);
                        // Only reachable on overflow (the host sizes the stack for the tree): every
                        // slot below CAP was written by this push, so popping stays defined.
                        if (sp > CAP) sp = CAP;
                    }
                }
            }
        }
        // Leaf phase: the lanes holding a leaf test its primitive records together, one record
        // per step, in record order (a lane keeps its place in an unfinished leaf for the next
        // round and walks no node before finishing it). Two steps per iteration, so that the
        // record under test and the one loading swap roles without register copies.
        while (true) {
            if (!leaf_go()) break;
            leaf_step<ANY, STATS, TRANSP>(r, prims, quads, transp, cur, alt, li, sp, bt, bp, bu, bv, occluded, gup, n_tris,
                                          n_quads);
            if (!leaf_go()) {
                cur = alt;
                break;
            }
            leaf_step<ANY, STATS, TRANSP>(r, prims, quads, transp, alt, cur, li, sp, bt, bp, bu, bv, occluded, gup, n_tris,
                                          n_quads);
        }
        if (occluded) sp = 0;
        if (TRANSP && ray_id != 0xffffffffu && sp == 0 && li == 0xffffffffu) {
            // Scene::GetAttenuation (scene.cpp:192-208) around Accelerator::GetAttenuation
            // (accelerator.cpp:26-68): blocked -> black; no hit -> the attenuation so far;
            // else multiply by the hit material's transparency (black without a transparent node)
            // and walk on from the hit point with (0.001, tmax - t).
            bool done = true;
            if (occluded) {
                att_r = att_g = att_b = 0.0f;
            } else if (bp != 0xffffffffu) {
                const float4 tr = transp ? reinterpret_cast<const float4*>(transp)[bp] : make_float4(0.f, 0.f, 0.f, 0.f);
                const bool has = tr.w != 0.0f;
                const float ar = has ? tr.x : 0.0f, ag = has ? tr.y : 0.0f, ab = has ? tr.z : 0.0f;
                if (ar == 0.0f && ag == 0.0f && ab == 0.0f) {
                    att_r = ar;
                    att_g = ag;
                    att_b = ab;
                } else {
                    att_r *= ar;
                    att_g *= ag;
                    att_b *= ab;
                    steps++;
                    if (!(att_r == 0.0f && att_g == 0.0f && att_b == 0.0f)) {
                        if (steps >= kMaxAttenuationSteps) {
                            atomicOr(a.error, kFlagAttenCap);
                        } else {
                            r.ox = r.ox + bt * r.dx;  // ray(t) = o + t * d
                            r.oy = r.oy + bt * r.dy;
                            r.oz = r.oz + bt * r.dz;
                            r.tmin = 0.001f;
                            r.tmax = r.tmax - bt;
                            ray_setup(r);
                            const TreeView& tv = MULTI ? a.trees[scene] : a.tree;
                            r.ieee = ieee_slabs_ok(r, tv.bbox);
                            const float rmin = tv.valid ? root_slab(r, tv.bbox) : -1.0f;
                            if (rmin >= 0.0f) {
                                push(0, tv.root, rmin);
                                sp = 1;
                                bt = FLT_MAX;
                                bp = 0xffffffffu;
                                bu = bv = 0.f;
                                gup = false;
                                done = false;
                            }
                        }
                    }
                }
            }
            if (done) {
                reinterpret_cast<float4*>(a.out)[(size_t)ray_id * ostride] = make_float4(att_r, att_g, att_b, (float)steps);
                ray_id = 0xffffffffu;
            }
        }
        if (!TRANSP && ray_id != 0xffffffffu && sp == 0 && li == 0xffffffffu) {
            if (ANY) {
                if (!a.occ_slot) {
                    static_cast<uint8_t*>(a.out)[ray_id] = occluded ? 1 : 0;
                } else if (occluded) {  // the shadow ray's contribution is cleared here (w kept)
                    const uint32_t slot = a.occ_slot[ray_id];
                    float* c = reinterpret_cast<float*>(a.occ_rec + 4 * (size_t)(slot >> 1) + 2 + (slot & 1u));
                    c[0] = 0.f;
                    c[1] = 0.f;
                    c[2] = 0.f;
                }
            } else {
                reinterpret_cast<float4*>(a.out)[ray_id] = make_float4(bt, __uint_as_float(bp), bu, bv);
            }
            ray_id = 0xffffffffu;
        }
        occluded = false;
    }

    if (overflow) atomicOr(a.error, kFlagStackOverflow);
    if (STATS) {
        unsigned long long v[5] = {n_nodes, n_leaves, n_tris, n_quads, n_wide};
#pragma unroll
        for (int k = 0; k < 5; k++) {
            unsigned long long x = v[k];
            for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
            if (lane == 0 && x) atomicAdd(a.stats + k, x);
        }
    }
}

__global__ void k_reset(uint32_t* counter) {
    if (threadIdx.x == 0) *counter = 0;
}

// Moves the error flags into flags[1] and clears them in one atomic step, so a flag raised by a
// kernel running meanwhile on another stream is either taken now or left for the next check.
__global__ void k_take_flags(uint32_t* flags) {
    if (threadIdx.x == 0) flags[1] = atomicExch(flags, 0u);
}

hipError_t take_flags(uint32_t* d_flags, hipStream_t stream, uint32_t* out) {
    *out = 0;
    hipLaunchKernelGGL(k_take_flags, dim3(1), dim3(64), 0, stream, d_flags);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_flags + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    return e;
}

template <int W, bool ANY, bool STATS, int SPILL, bool MULTI, int LDS, bool TRANSP = false>
static hipError_t launch_one(const TraceArgs& a, hipStream_t stream, int num_cus) {
    auto kern = k_trace<W, ANY, STATS, SPILL, MULTI, LDS, TRANSP>;
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    if (a.blocks_per_cu && (int)a.blocks_per_cu < per_cu) per_cu = (int)a.blocks_per_cu;
    const uint64_t want = ((uint64_t)a.n + kBlock - 1) / kBlock;
    uint64_t grid = (uint64_t)num_cus * (uint64_t)per_cu;
    if (want < grid) grid = want;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, stream, a.counter);
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(kBlock), 0, stream, a);
    return hipGetLastError();
}

template <int W, bool ANY, bool STATS, bool MULTI, int LDS>
static hipError_t launch_spill(const TraceArgs& a, uint32_t need, hipStream_t stream, int num_cus) {
    if (need <= (uint32_t)(LDS + 48)) return launch_one<W, ANY, STATS, 48, MULTI, LDS>(a, stream, num_cus);
    if (need <= (uint32_t)(LDS + 240)) return launch_one<W, ANY, STATS, 240, MULTI, LDS>(a, stream, num_cus);
    return hipErrorInvalidValue;
}

template <int W, bool ANY, bool STATS, bool MULTI>
static hipError_t launch_lds(const TraceArgs& a, uint32_t need, hipStream_t stream, int num_cus) {
    if (!STATS && a.lds_entries == 8) return launch_spill<W, ANY, STATS, MULTI, 8>(a, need, stream, num_cus);
    return launch_spill<W, ANY, STATS, MULTI, 16>(a, need, stream, num_cus);
}

template <int W, bool MULTI>
static hipError_t launch_w(const TraceArgs& a, bool any, bool stats, hipStream_t stream, int num_cus) {
    const uint32_t need = MULTI ? a.max_stack : a.tree.max_stack;
    if (any) return stats ? launch_lds<W, true, true, MULTI>(a, need, stream, num_cus)
                          : launch_lds<W, true, false, MULTI>(a, need, stream, num_cus);
    return stats ? launch_lds<W, false, true, MULTI>(a, need, stream, num_cus)
                 : launch_lds<W, false, false, MULTI>(a, need, stream, num_cus);
}

hipError_t launch_trace(const TraceArgs& a, bool any, bool stats, hipStream_t stream, int num_cus) {
    if (a.n == 0) return hipSuccess;
    const bool multi = a.trees != nullptr;
    const uint32_t width = multi ? a.width : a.tree.width;
    if (width == 8) return multi ? launch_w<8, true>(a, any, stats, stream, num_cus)
                                 : launch_w<8, false>(a, any, stats, stream, num_cus);
    return multi ? launch_w<4, true>(a, any, stats, stream, num_cus) : launch_w<4, false>(a, any, stats, stream, num_cus);
}

template <int W, bool MULTI>
static hipError_t launch_att_w(const TraceArgs& a, hipStream_t stream, int num_cus) {
    const uint32_t need = MULTI ? a.max_stack : a.tree.max_stack;
    if (need <= 16u + 48u) return launch_one<W, false, false, 48, MULTI, 16, true>(a, stream, num_cus);
    if (need <= 16u + 240u) return launch_one<W, false, false, 240, MULTI, 16, true>(a, stream, num_cus);
    return hipErrorInvalidValue;
}

hipError_t launch_attenuation(const TraceArgs& a, hipStream_t stream, int num_cus) {
    if (a.n == 0) return hipSuccess;
    if (a.trees) {
        if (!a.transps) return hipErrorInvalidValue;
        return a.width == 8 ? launch_att_w<8, true>(a, stream, num_cus) : launch_att_w<4, true>(a, stream, num_cus);
    }
    if (!a.transp) return hipErrorInvalidValue;
    return a.tree.width == 8 ? launch_att_w<8, false>(a, stream, num_cus) : launch_att_w<4, false>(a, stream, num_cus);
}

uint32_t trace_max_stack_supported() { return 8 + 240; }

}  // namespace sortgpu
