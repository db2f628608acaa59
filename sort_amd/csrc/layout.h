// Device record layout of a flattened Fast-BVH, shared by the host flattener and the HIP
// kernels.
//
// Nodes: one record per inner node, SoA over the W children so that one lane reads a whole
// node with 16-byte loads: six bound planes (lo_x, hi_x, lo_y, hi_y, lo_z, hi_z; W floats
// each) then W child references. W = 8: 192 + 32 = 224 B, padded to 256 B (two 128-byte
// lines). W = 4: 96 + 16 = 112 B, padded to 128 B.
//
// Child reference (uint32):
//   0xFFFFFFFF            empty slot (the reference's BBox8 m_mask lane = 0)
//   bit31 = 0             inner child: node index
//   bit31 = 1             leaf: index of its first primitive record
//
// Leaf records: 80 bytes (five 16-byte loads), in tree order. A leaf's records are contiguous:
// its triangles in build order, two per record (A then B; the last record of an odd count holds
// one), then its quads, one per record (the reference tests a leaf's Triangle8 groups first and
// its other_list after, fast_bvh.hpp:349-406). A leaf walk loads one record per step, so a pair
// record halves the steps of a triangle walk. Flags live in the top bits of the two primitive
// ids, which limits a tree to 2^30 - 1 primitives.
#pragma once
#include <stdint.h>

namespace sortgpu {

constexpr uint32_t REF_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t REF_LEAF = 0x80000000u;

constexpr uint32_t PR_ID = 0x3FFFFFFFu;    // primitive id bits of id_a / id_b
constexpr uint32_t PR_NONE = 0x3FFFFFFFu;  // id_b: the record holds triangle A only
constexpr uint32_t PR_FLAG = 0x80000000u;  // id_a: the record is a quad; id_b: last record of the leaf
constexpr uint32_t PR_GEND = 0x40000000u;  // id_a / id_b: that triangle ends a group of W (the
                                           // reference's Simd_Triangle)

template <int W> struct alignas(W == 8 ? 256 : 128) NodeRec {
    float lo_x[W], hi_x[W], lo_y[W], hi_y[W], lo_z[W], hi_z[W];
    uint32_t child[W];
    uint32_t pad[W == 8 ? 8 : 4];
};
static_assert(sizeof(NodeRec<8>) == 256, "Node8 must be 256 B");
static_assert(sizeof(NodeRec<4>) == 128, "Node4 must be 128 B");

// A triangle's vertices as the watertight test reads them.
struct Tri {
    float p0x, p0y, p0z, p1x, p1y, p1z, p2x, p2y, p2z;
};

struct alignas(16) PrimRec {
    Tri a, b;        // triangles A and B; a quad record keeps its quad index in a.p0x's bits
    uint32_t id_a;   // A's primitive id | PR_GEND; a quad: its primitive id | PR_FLAG
    uint32_t id_b;   // B's primitive id (PR_NONE: no B) | PR_GEND; | PR_FLAG on the leaf's last record
};
static_assert(sizeof(PrimRec) == 80, "PrimRec must be 80 B");

// Quad light record: inverse transform (world -> light), half sizes, primitive id.
struct alignas(16) QuadRec {
    float inv[16];
    float half_x, half_y;
    uint32_t prim;
    uint32_t pad;
};
static_assert(sizeof(QuadRec) == 80, "QuadRec must be 80 B");

// Everything a traversal kernel needs about one tree.
struct TreeView {
    const void* nodes;      // NodeRec<W>*
    const PrimRec* prims;
    const QuadRec* quads;
    uint32_t root;          // child reference of the root (leaf when the tree is one leaf)
    uint32_t valid;         // 0: every query misses
    uint32_t max_stack;     // (W-1)*(depth-1)+1
    uint32_t width;
    float bbox[6];          // enlarged scene box (root slab test)
};

}  // namespace sortgpu
