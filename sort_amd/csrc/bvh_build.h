// Host-side SAH Fast-BVH builder (8-wide OBVH / 4-wide QBVH) and flattener.
//
// Split decisions follow SORT's Fbvh::splitNode / pickBestSplit (accel/fast_bvh.hpp:138-211,
// accel/bvh_utils.h:56-116) so the tree — node structure, primitive order, child order and
// bound planes — is the reference's tree. Independent subtrees are built on a thread pool;
// every range is still processed by one thread in the reference's order, so the output does
// not depend on the thread count.
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

#include "layout.h"

namespace sortgpu {

struct QuadIn {
    float mat[16];
    float inv[16];
    float size_x, size_y;
};

struct BuildInput {
    const float* pos = nullptr;
    uint32_t nverts = 0;
    const uint32_t* idx = nullptr;
    uint32_t ntris = 0;
    const QuadIn* quads = nullptr;
    uint32_t nquads = 0;
    uint32_t width = 8;
    uint32_t max_depth = 28;
    uint32_t max_leaf = 16;
    uint32_t threads = 0;
};

struct BuiltTree {
    uint32_t width = 8;
    uint32_t depth = 0;       // deepest leaf level, root = 1 (Fbvh::m_depth)
    uint32_t n_inner = 0;     // inner nodes (flattened node records)
    uint32_t n_leaves = 0;
    uint32_t nprims = 0;
    uint32_t root_ref = REF_EMPTY;
    uint32_t max_stack = 0;
    bool valid = false;
    float bbox[6] = {0, 0, 0, 0, 0, 0};
    double build_ms = 0.0;
    std::vector<uint32_t> prim_order;       // m_bvhpri order
    std::vector<uint32_t> dump;             // pre-order (child_cnt, pri_offset, pri_cnt)
    std::vector<unsigned char> node_bytes;  // n_inner NodeRec<width>
    std::vector<PrimRec> prims;             // leaf records (+1 zero record)
    std::vector<QuadRec> quads;
};

// Returns false and sets err on invalid input.
bool build_tree(const BuildInput& in, BuiltTree& out, std::string& err);

}  // namespace sortgpu
