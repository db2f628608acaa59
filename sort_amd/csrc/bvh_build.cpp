// Host SAH Fast-BVH builder + flattener. See bvh_build.h.
//
// Arithmetic must match SORT's build bit for bit (it decides the tree): this file is compiled
// with -ffp-contract=off and every expression keeps the reference's operation order.
//   per-primitive box           shape/triangle.cpp:116-136, shape/quad.cpp:96-108
//   scene box (+0.1% per side)  core/scene.cpp:145-154
//   centroid                    accel/bvh_utils.h:35-38
//   SAH bins                    accel/bvh_utils.h:56-116
//   8-way BFS split             accel/fast_bvh.hpp:138-211
//   leaf / child boxes          accel/fast_bvh.hpp:100-107,213-302
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <queue>
#include <thread>

namespace sortgpu {
namespace {

// libstdc++'s std::min / std::max operand order (decides -0/+0 in unions).
inline float smin(float a, float b) { return (b < a) ? b : a; }
inline float smax(float a, float b) { return (a < b) ? b : a; }

struct Box {
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    inline void add_point(float x, float y, float z) {  // BBox::Union(Point)
        const float p[3] = {x, y, z};
        for (int i = 0; i < 3; i++) {
            if (p[i] < mn[i]) mn[i] = p[i];
            if (p[i] > mx[i]) mx[i] = p[i];
        }
    }
    inline void add_box(const float* bmn, const float* bmx) {  // BBox::Union(BBox)
        for (int i = 0; i < 3; i++) {
            mn[i] = smin(mn[i], bmn[i]);
            mx[i] = smax(mx[i], bmx[i]);
        }
    }
    inline void add_box(const Box& b) { add_box(b.mn, b.mx); }
    inline float half_area() const {
        const float ox = mx[0] - mn[0], oy = mx[1] - mn[1], oz = mx[2] - mn[2];
        return ox * oy + oy * oz + oz * ox;
    }
    inline unsigned max_axis() const {
        const float ox = mx[0] - mn[0], oy = mx[1] - mn[1], oz = mx[2] - mn[2];
        if (ox > oy && ox > oz) return 0;
        if (oy > ox && oy > oz) return 1;
        return 2;
    }
};
inline Box union_of(const Box& a, const Box& b) {  // free Union(BBox, BBox), bbox.h:115-122
    Box r;
    for (int i = 0; i < 3; i++) {
        r.mn[i] = smin(a.mn[i], b.mn[i]);
        r.mx[i] = smax(a.mx[i], b.mx[i]);
    }
    return r;
}

// Build primitive: bounds, centroid and id together so that binning and partitioning stream
// one array (40 B per primitive).
struct BPrim {
    float mn[3];
    float mx[3];
    float c[3];
    uint32_t id;
};

struct BNode {
    uint32_t off = 0, cnt = 0, child_cnt = 0;
    uint32_t rec = 0xFFFFFFFFu;  // flattened leaf: index of its first leaf record
    std::unique_ptr<BNode> child[8];
    BNode(uint32_t o, uint32_t c) : off(o), cnt(c) {}
};

// Minimal task pool: tasks may submit more tasks; wait() returns when all are done.
class Pool {
  public:
    explicit Pool(unsigned n) {
        for (unsigned i = 0; i < n; i++) th_.emplace_back([this] { loop(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> l(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    bool parallel() const { return !th_.empty(); }
    void submit(std::function<void()> f) {
        pending_.fetch_add(1);
        {
            std::lock_guard<std::mutex> l(m_);
            q_.push_back(std::move(f));
        }
        cv_.notify_one();
    }
    void wait() {
        // The waiting thread helps until every task is done.
        while (pending_.load() > 0) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> l(m_);
                if (!q_.empty()) {
                    f = std::move(q_.back());
                    q_.pop_back();
                }
            }
            if (f) {
                f();
                pending_.fetch_sub(1);
            } else {
                std::this_thread::yield();
            }
        }
    }

  private:
    void loop() {
        while (true) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> l(m_);
                cv_.wait(l, [this] { return stop_ || !q_.empty(); });
                if (stop_ && q_.empty()) return;
                f = std::move(q_.back());
                q_.pop_back();
            }
            f();
            pending_.fetch_sub(1);
        }
    }
    std::vector<std::thread> th_;
    std::deque<std::function<void()>> q_;
    std::mutex m_;
    std::condition_variable cv_;
    std::atomic<int64_t> pending_{0};
    bool stop_ = false;
};

struct Builder {
    std::vector<BPrim> P;
    unsigned W = 8, max_leaf = 16, max_depth = 28;
    std::atomic<uint32_t> depth{0}, n_inner{0}, n_leaves{0};
    Pool* pool = nullptr;
    static constexpr uint32_t kParallelMin = 8192;  // ranges below this recurse inline

    Box range_box(uint32_t off, uint32_t cnt) const {  // calcBoundingBox, fast_bvh.hpp:100-107
        Box b;
        for (uint32_t i = off; i < off + cnt; i++) b.add_box(P[i].mn, P[i].mx);
        return b;
    }

    // accel/bvh_utils.h:69-116
    float pick_best_split(unsigned& axis, float& split_pos, const Box& node_box, uint32_t start,
                          uint32_t end) const {
        constexpr unsigned N = 16;
        constexpr float INV = 1.0f / (float)N;
        Box inner;
        for (uint32_t i = start; i < end; i++) inner.add_point(P[i].c[0], P[i].c[1], P[i].c[2]);
        const uint32_t primitive_num = end - start;
        axis = inner.max_axis();
        float min_sah = FLT_MAX;
        unsigned bin[N] = {0};
        Box bbox[N];
        Box rbox[N - 1];
        const float split_start = inner.mn[axis];
        const float split_delta = (inner.mx[axis] - inner.mn[axis]) * INV;
        if (split_delta == 0.0f) return FLT_MAX;
        const float inv_split_delta = 1.0f / split_delta;
        for (uint32_t i = start; i < end; i++) {
            int index = (int)((P[i].c[axis] - split_start) * inv_split_delta);
            index = std::min(index, (int)(N - 1));
            ++bin[index];
            bbox[index].add_box(P[i].mn, P[i].mx);
        }
        rbox[N - 2].add_box(bbox[N - 1]);
        for (int i = N - 3; i >= 0; i--) rbox[i] = union_of(rbox[i + 1], bbox[i + 1]);
        unsigned left = bin[0];
        Box lbox = bbox[0];
        float pos = split_delta + split_start;
        const float node_half = node_box.half_area();
        for (unsigned i = 0; i < N - 1; i++) {
            const unsigned right = primitive_num - left;
            const float sah = (left * lbox.half_area() + right * rbox[i].half_area()) / node_half;
            if (sah < min_sah) {
                min_sah = sah;
                split_pos = pos;
            }
            left += bin[i + 1];
            lbox.add_box(bbox[i + 1]);
            pos += split_delta;
        }
        return min_sah;
    }

    void make_leaf(BNode* node, unsigned d) {
        node->child_cnt = 0;
        uint32_t cur = depth.load();
        while (d > cur && !depth.compare_exchange_weak(cur, d)) {
        }
        n_leaves.fetch_add(1);
    }

    // accel/fast_bvh.hpp:138-211
    void split(BNode* node, const Box& node_box, unsigned d) {
        const uint32_t start = node->off, end = start + node->cnt;
        if (node->cnt <= max_leaf || d == max_depth) {
            make_leaf(node, d);
            return;
        }
        std::queue<std::pair<uint32_t, uint32_t>> to_split, done;
        to_split.push({start, end});
        while (!to_split.empty() && to_split.size() + done.size() < W) {
            const auto cur = to_split.front();
            to_split.pop();
            const uint32_t s = cur.first, e = cur.second, cnt = e - s;
            unsigned axis;
            float pos;
            const float sah = pick_best_split(axis, pos, node_box, s, e);
            if (sah >= cnt || cnt <= max_leaf) {
                done.push(cur);
            } else {
                const auto pred = [pos, axis](const BPrim& p) { return p.c[axis] < pos; };
                BPrim* middle = std::partition(P.data() + s, P.data() + e, pred);
                const uint32_t mid = (uint32_t)(middle - P.data());
                if (mid == s || mid == e) {
                    done.push(cur);
                } else {
                    to_split.push({s, mid});
                    to_split.push({mid, e});
                }
            }
        }
        if (to_split.size() + done.size() == 1) {
            make_leaf(node, d);
            return;
        }
        while (!to_split.empty()) {
            const auto c = to_split.front();
            to_split.pop();
            node->child[node->child_cnt++].reset(new BNode(c.first, c.second - c.first));
        }
        while (!done.empty()) {
            const auto c = done.front();
            done.pop();
            node->child[node->child_cnt++].reset(new BNode(c.first, c.second - c.first));
        }
        n_inner.fetch_add(1);
        for (uint32_t j = 0; j < node->child_cnt; ++j) {
            BNode* ch = node->child[j].get();
            if (pool && pool->parallel() && ch->cnt >= kParallelMin) {
                pool->submit([this, ch, d] {
                    const Box b = range_box(ch->off, ch->cnt);
                    split(ch, b, d + 1);
                });
            } else {
                const Box b = range_box(ch->off, ch->cnt);
                split(ch, b, d + 1);
            }
        }
    }
};

template <int W>
struct Flattener {
    const Builder& B;
    const BuildInput& in;
    std::vector<NodeRec<W>> nodes;
    std::vector<PrimRec> prims;
    std::vector<uint32_t> dump;
    uint32_t next_rec = 0;  // leaf records assigned so far

    Flattener(const Builder& b, const BuildInput& i) : B(b), in(i) {}

    // The leaf's record range, assigned when its reference is first written (its parent's
    // record is filled before the leaf's records are), in that order.
    uint32_t leaf_ref(BNode* n) {
        if (n->rec == 0xFFFFFFFFu) {
            uint32_t tris = 0;
            for (uint32_t i = n->off; i < n->off + n->cnt; i++) tris += B.P[i].id < in.ntris ? 1u : 0u;
            n->rec = next_rec;
            next_rec += (tris + 1) / 2 + (n->cnt - tris);
        }
        return REF_LEAF | n->rec;
    }

    // Leaf records: triangles (build order) two per record, then quads one per record, the last
    // record flagged. The triangles form groups of W in that order, the last one possibly partial
    // (the reference packs a leaf's triangles into Simd_Triangle groups this way,
    // fast_bvh.hpp:226-251); the last triangle of each group carries PR_GEND.
    void emit_leaf(BNode* n) {
        const uint32_t first = leaf_ref(n) & ~REF_LEAF;
        uint32_t w = first;
        uint32_t j = 0;  // triangles emitted so far in this leaf
        uint32_t tris = 0;
        for (uint32_t i = n->off; i < n->off + n->cnt; i++) tris += B.P[i].id < in.ntris ? 1u : 0u;
        for (uint32_t i = n->off; i < n->off + n->cnt; i++) {
            const uint32_t id = B.P[i].id;
            if (id >= in.ntris) continue;
            PrimRec& r = prims[w];
            Tri& x = (j & 1u) ? r.b : r.a;
            const uint32_t* t = in.idx + 3 * (size_t)id;
            const float* a = in.pos + 3 * (size_t)t[0];
            const float* b = in.pos + 3 * (size_t)t[1];
            const float* c = in.pos + 3 * (size_t)t[2];
            x.p0x = a[0]; x.p0y = a[1]; x.p0z = a[2];
            x.p1x = b[0]; x.p1y = b[1]; x.p1z = b[2];
            x.p2x = c[0]; x.p2y = c[1]; x.p2z = c[2];
            const uint32_t gend = ((j + 1) % W == 0 || j + 1 == tris) ? PR_GEND : 0u;
            if (j & 1u) {
                r.id_b = id | gend;
                w++;
            } else {
                r.id_a = id | gend;
                r.id_b = PR_NONE;
            }
            j++;
        }
        if (j & 1u) w++;
        for (uint32_t i = n->off; i < n->off + n->cnt; i++) {
            const uint32_t id = B.P[i].id;
            if (id < in.ntris) continue;
            PrimRec& r = prims[w++];
            std::memset(&r, 0, sizeof(r));
            const uint32_t q = id - in.ntris;
            std::memcpy(&r.a.p0x, &q, sizeof(q));
            r.id_a = id | PR_FLAG;
            r.id_b = PR_NONE;
        }
        prims[w - 1].id_b |= PR_FLAG;
    }

    // Children of a node get consecutive node indices.
    void emit_inner(const BNode* n, uint32_t self) {
        uint32_t child_idx[8];
        for (uint32_t j = 0; j < n->child_cnt; j++) {
            BNode* ch = n->child[j].get();
            if (ch->child_cnt) {
                child_idx[j] = (uint32_t)nodes.size();
                nodes.emplace_back();
            }
        }
        NodeRec<W>& r = nodes[self];
        std::memset(&r, 0, sizeof(r));
        for (int j = 0; j < W; j++) {
            // calcBoundingBoxSIMD (fast_bvh.hpp:273-302): empty slots keep the invalid box.
            Box b;
            if ((uint32_t)j < n->child_cnt) b = B.range_box(n->child[j]->off, n->child[j]->cnt);
            r.lo_x[j] = b.mn[0]; r.hi_x[j] = b.mx[0];
            r.lo_y[j] = b.mn[1]; r.hi_y[j] = b.mx[1];
            r.lo_z[j] = b.mn[2]; r.hi_z[j] = b.mx[2];
            if ((uint32_t)j >= n->child_cnt) {
                r.child[j] = REF_EMPTY;
            } else {
                BNode* ch = n->child[j].get();
                r.child[j] = ch->child_cnt ? child_idx[j] : leaf_ref(ch);
            }
        }
        for (uint32_t j = 0; j < n->child_cnt; j++) {
            BNode* ch = n->child[j].get();
            if (ch->child_cnt) emit_inner(ch, child_idx[j]);
            else emit_leaf(ch);
        }
    }

    void dump_node(const BNode* n) {
        dump.push_back(n->child_cnt);
        dump.push_back(n->off);
        dump.push_back(n->cnt);
        for (uint32_t j = 0; j < n->child_cnt; j++) dump_node(n->child[j].get());
    }

    uint32_t run(BNode* root, uint32_t nprims) {
        // at most one record per primitive, plus a zero record after the last
        prims.assign(nprims + 1, PrimRec{});
        dump_node(root);
        uint32_t ref = 0;
        if (!root->child_cnt) {
            emit_leaf(root);
            ref = leaf_ref(root);
        } else {
            nodes.reserve(B.n_inner.load());
            nodes.emplace_back();
            emit_inner(root, 0);
        }
        prims.resize((size_t)next_rec + 1);
        return ref;
    }
};

}  // namespace

bool build_tree(const BuildInput& in, BuiltTree& out, std::string& err) {
    const auto t0 = std::chrono::steady_clock::now();
    if (in.width != 4 && in.width != 8) {
        err = "width must be 4 or 8";
        return false;
    }
    if (in.max_depth < 1 || in.max_leaf < 1) {
        err = "max_depth and max_leaf must be >= 1";
        return false;
    }
    if ((in.ntris && (!in.pos || !in.idx)) || (in.nquads && !in.quads)) {
        err = "null geometry pointer";
        return false;
    }
    const uint64_t n64 = (uint64_t)in.ntris + in.nquads;
    if (n64 > PR_ID) {  // primitive ids share their word with two record flags
        err = "too many primitives (limit 2^30-1)";
        return false;
    }
    for (uint64_t i = 0; i < 3ull * in.ntris; i++) {
        if (in.idx[i] >= in.nverts) {
            err = "vertex index out of range";
            return false;
        }
    }
    const uint32_t n = (uint32_t)n64;
    out = BuiltTree();
    out.width = in.width;
    out.nprims = n;

    Builder B;
    B.W = in.width;
    B.max_leaf = in.max_leaf;
    B.max_depth = in.max_depth;
    B.P.resize(n);

    unsigned nthreads = in.threads ? in.threads : std::max(1u, std::thread::hardware_concurrency());
    nthreads = std::min(nthreads, 64u);

    // Primitive boxes and centroids, in primitive-id order.
    auto prim_boxes = [&](uint32_t lo, uint32_t hi) {
        for (uint32_t t = lo; t < hi; t++) {
            Box b;
            if (t < in.ntris) {
                const uint32_t* v = in.idx + 3 * (size_t)t;
                for (int k = 0; k < 3; k++) {
                    const float* p = in.pos + 3 * (size_t)v[k];
                    b.add_point(p[0], p[1], p[2]);
                }
            } else {
                const QuadIn& q = in.quads[t - in.ntris];
                const float hx = q.size_x * 0.5f, hy = q.size_y * 0.5f;
                const float corners[4][3] = {{hx, 0.f, hy}, {hx, 0.f, -hy}, {-hx, 0.f, hy}, {-hx, 0.f, -hy}};
                for (int k = 0; k < 4; k++) {  // Matrix::TransformPoint, math/matrix.cpp:70-82
                    const float* m = q.mat;
                    const float* p = corners[k];
                    const float x = p[0] * m[0] + p[1] * m[1] + p[2] * m[2] + m[3];
                    const float y = p[0] * m[4] + p[1] * m[5] + p[2] * m[6] + m[7];
                    const float z = p[0] * m[8] + p[1] * m[9] + p[2] * m[10] + m[11];
                    const float w = p[0] * m[12] + p[1] * m[13] + p[2] * m[14] + m[15];
                    if (w == 1.0f) b.add_point(x, y, z);
                    else b.add_point(x / w, y / w, z / w);
                }
            }
            BPrim& bp = B.P[t];
            for (int i = 0; i < 3; i++) {
                bp.mn[i] = b.mn[i];
                bp.mx[i] = b.mx[i];
                bp.c[i] = (b.mx[i] + b.mn[i]) * 0.5f;
            }
            bp.id = t;
        }
    };
    {
        std::vector<std::thread> th;
        const uint32_t chunk = (n + nthreads - 1) / std::max(1u, nthreads);
        for (unsigned w = 0; w < nthreads && (uint64_t)w * chunk < n; w++)
            th.emplace_back(prim_boxes, w * chunk, std::min<uint32_t>(n, (w + 1) * chunk));
        for (auto& t : th) t.join();
    }

    Box sb;
    for (uint32_t i = 0; i < n; i++) sb.add_box(B.P[i].mn, B.P[i].mx);
    {
        const float dx = (sb.mx[0] - sb.mn[0]) * 0.001f;
        const float dy = (sb.mx[1] - sb.mn[1]) * 0.001f;
        const float dz = (sb.mx[2] - sb.mn[2]) * 0.001f;
        sb.mn[0] -= dx; sb.mn[1] -= dy; sb.mn[2] -= dz;
        sb.mx[0] += dx; sb.mx[1] += dy; sb.mx[2] += dz;
    }
    for (int i = 0; i < 3; i++) {
        out.bbox[i] = sb.mn[i];
        out.bbox[3 + i] = sb.mx[i];
    }

    for (uint32_t q = 0; q < in.nquads; q++) {
        QuadRec r;
        std::memset(&r, 0, sizeof(r));
        std::memcpy(r.inv, in.quads[q].inv, sizeof(r.inv));
        r.half_x = in.quads[q].size_x * 0.5f;
        r.half_y = in.quads[q].size_y * 0.5f;
        r.prim = in.ntris + q;
        out.quads.push_back(r);
    }

    if (n == 0) {
        out.valid = false;
        out.prims.assign(1, PrimRec{});
        out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return true;
    }

    std::unique_ptr<BNode> root(new BNode(0, n));
    {
        std::unique_ptr<Pool> pool;
        if (nthreads > 1) pool.reset(new Pool(nthreads - 1));
        B.pool = pool.get();
        B.split(root.get(), sb, 1u);
        if (pool) pool->wait();
    }

    out.valid = true;
    out.depth = B.depth.load();
    out.n_inner = B.n_inner.load();
    out.n_leaves = B.n_leaves.load();
    out.max_stack = (in.width - 1) * (out.depth - 1) + 1;
    out.prim_order.resize(n);
    for (uint32_t i = 0; i < n; i++) out.prim_order[i] = B.P[i].id;

    if (in.width == 8) {
        Flattener<8> F(B, in);
        out.root_ref = F.run(root.get(), n);
        out.node_bytes.resize(F.nodes.size() * sizeof(NodeRec<8>));
        if (!F.nodes.empty()) std::memcpy(out.node_bytes.data(), F.nodes.data(), out.node_bytes.size());
        out.prims = std::move(F.prims);
        out.dump = std::move(F.dump);
    } else {
        Flattener<4> F(B, in);
        out.root_ref = F.run(root.get(), n);
        out.node_bytes.resize(F.nodes.size() * sizeof(NodeRec<4>));
        if (!F.nodes.empty()) std::memcpy(out.node_bytes.data(), F.nodes.data(), out.node_bytes.size());
        out.prims = std::move(F.prims);
        out.dump = std::move(F.dump);
    }
    out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

}  // namespace sortgpu
