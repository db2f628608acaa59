// Devices of the serving engine: the HIP device (scenes, films and wave workspaces in HBM, waves
// on a few non-blocking streams) and a host-only stand-in with the same budget arithmetic,
// whose films hold a hash of each path key, for tests of scheduling and failure handling.
#include "serve_device.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>

#include "pt_host.h"

namespace sortgpu {
namespace serve {

namespace {

std::string hip_msg(const char* what, hipError_t e) { return std::string(what) + ": " + hipGetErrorString(e); }

int hip_rc(hipError_t e) { return e == hipErrorOutOfMemory ? SORTGPU_ERR_OOM : SORTGPU_ERR_HIP; }

// Film resolve: per-pixel mean of the valid samples, the same float operations as the device
// resolve of the single-render path.
void resolve(const float* acc, uint32_t npix, std::vector<float>& rgb) {
    rgb.resize((size_t)npix * 3);
    for (uint32_t i = 0; i < npix; i++) {
        const float* a = acc + 4 * (size_t)i;
        const float inv = a[3] > 0.0f ? 1.0f / a[3] : 0.0f;
        rgb[3 * (size_t)i] = a[0] * inv;
        rgb[3 * (size_t)i + 1] = a[1] * inv;
        rgb[3 * (size_t)i + 2] = a[2] * inv;
    }
}

// Waves alternate over `nstreams` streams, each with its own path workspace, so the latency-bound
// last bounces of one wave overlap the first bounces of the next. Film updates stay in wave order
// (each waits for the previous wave's), scene uploads run on a stream of their own and every
// wave waits for the latest upload.
class HipDevice : public Device {
  public:
    static constexpr int kMaxStreams = 4;

    HipDevice(sortgpu_ctx* ctx, uint32_t wave_paths, uint32_t max_scenes, uint32_t nstreams, bool graphs)
        : ctx_(ctx), cap_(wave_paths), max_scenes_(max_scenes), ns_(std::max(1u, std::min(nstreams, 4u))),
          graphs_(graphs) {}

    int init(std::string& err) {
        int rc = ctx_bind(ctx_);
        if (rc) return (err = "cannot bind the device", rc);
        hipError_t e;
        for (uint32_t i = 0; i < ns_; i++) {
            if ((e = hipStreamCreateWithFlags(&s_[i], hipStreamNonBlocking)) != hipSuccess)
                return (err = hip_msg("hipStreamCreate", e), hip_rc(e));
            if ((rc = ws_[i].ensure(ctx_, cap_))) return (err = "path workspace allocation failed", rc);
            void* p = nullptr;
            if ((e = hipMalloc(&p, sizeof(WaveDesc))) != hipSuccess)
                return (err = hip_msg("hipMalloc(wave descriptor)", e), hip_rc(e));
            d_desc_[i] = static_cast<WaveDesc*>(p);
        }
        if ((e = hipStreamCreateWithFlags(&s_up_, hipStreamNonBlocking)) != hipSuccess)
            return (err = hip_msg("hipStreamCreate", e), hip_rc(e));
        if ((e = hipEventCreateWithFlags(&ev_up_, hipEventDisableTiming)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&ev_film_, hipEventDisableTiming)) != hipSuccess)
            return (err = hip_msg("hipEventCreate", e), hip_rc(e));
        if ((rc = ring_.init(ctx_))) return (err = "request ring allocation failed", rc);
        const size_t sb = (size_t)max_scenes_ * sizeof(SceneDev), tb = (size_t)max_scenes_ * sizeof(TreeView);
        const size_t cb = (size_t)kCounterSlots * kCounterStride * sizeof(uint32_t);
        void* p = nullptr;
        if ((e = hipMalloc(&p, sb)) != hipSuccess) return (err = hip_msg("hipMalloc(scene table)", e), hip_rc(e));
        d_scenes_ = static_cast<SceneDev*>(p);
        if ((e = hipMalloc(&p, tb)) != hipSuccess) return (err = hip_msg("hipMalloc(tree table)", e), hip_rc(e));
        d_trees_ = static_cast<TreeView*>(p);
        const size_t pb = (size_t)max_scenes_ * sizeof(const float*);
        if ((e = hipMalloc(&p, pb)) != hipSuccess) return (err = hip_msg("hipMalloc(transparency table)", e), hip_rc(e));
        d_transps_ = static_cast<const float**>(p);
        if ((e = hipMalloc(&p, cb + 256)) != hipSuccess) return (err = hip_msg("hipMalloc(counters)", e), hip_rc(e));
        d_counters_ = static_cast<uint32_t*>(p);
        d_error_ = d_counters_ + cb / sizeof(uint32_t);
        if ((e = hipHostMalloc(&p, sb, hipHostMallocDefault)) != hipSuccess)
            return (err = hip_msg("hipHostMalloc(scene table)", e), hip_rc(e));
        h_scenes_ = static_cast<SceneDev*>(p);
        if ((e = hipHostMalloc(&p, tb, hipHostMallocDefault)) != hipSuccess)
            return (err = hip_msg("hipHostMalloc(tree table)", e), hip_rc(e));
        h_trees_ = static_cast<TreeView*>(p);
        if ((e = hipHostMalloc(&p, pb, hipHostMallocDefault)) != hipSuccess)
            return (err = hip_msg("hipHostMalloc(transparency table)", e), hip_rc(e));
        h_transps_ = static_cast<const float**>(p);
        std::memset(h_scenes_, 0, sb);
        std::memset(h_trees_, 0, tb);
        std::memset(h_transps_, 0, pb);
        if ((e = hipMemcpy(d_scenes_, h_scenes_, sb, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(d_trees_, h_trees_, tb, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(d_transps_, h_transps_, pb, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemset(d_counters_, 0, cb + 256)) != hipSuccess)
            return (err = hip_msg("table initialisation", e), hip_rc(e));
        for (uint32_t i = 0; i < ns_; i++)
            if ((e = pt_zero_f4(reinterpret_cast<float4*>(ws_[i].q.totals), 1, s_[i])) != hipSuccess ||
                (e = hipStreamSynchronize(s_[i])) != hipSuccess)
                return (err = hip_msg("workspace initialisation", e), hip_rc(e));
        fixed_ = sb + tb + pb + cb + 256 + (size_t)ReqRing::kSlots * sizeof(WaveDesc);
        for (uint32_t i = 0; i < ns_; i++) fixed_ += ws_[i].mem.bytes + sizeof(WaveDesc);
        return SORTGPU_OK;
    }

    ~HipDevice() override {
        ctx_bind(ctx_);
        sync_all();
        drop_graphs();
        for (uint32_t i = 0; i < ns_; i++)
            if (d_desc_[i]) hipFree(d_desc_[i]);
        for (SceneHost& sh : scenes_)
            if (sh.block) hipFree(sh.block);
        for (auto& kv : cache_) hipFree(kv.second);
        for (Ev& ev : evs_) spare_.push_back(ev.a), spare_.push_back(ev.b);
        for (hipEvent_t e : spare_) hipEventDestroy(e);
        ring_.release();
        for (uint32_t i = 0; i < ns_; i++) ws_[i].release();
        if (d_scenes_) hipFree(d_scenes_);
        if (d_trees_) hipFree(d_trees_);
        if (d_counters_) hipFree(d_counters_);
        if (h_scenes_) hipHostFree(h_scenes_);
        if (h_trees_) hipHostFree(h_trees_);
        if (d_transps_) hipFree(d_transps_);
        if (h_transps_) hipHostFree(h_transps_);
        if (ev_up_) hipEventDestroy(ev_up_);
        if (ev_film_) hipEventDestroy(ev_film_);
        for (uint32_t i = 0; i < ns_; i++)
            if (s_[i]) hipStreamDestroy(s_[i]);
        if (s_up_) hipStreamDestroy(s_up_);
    }

    size_t fixed_bytes() const override { return fixed_; }
    uint32_t wave_paths() const override { return cap_; }

    int add_scene(const sortgpu_accel* acc, const sortgpu_scene_desc* desc, uint32_t id, size_t* bytes,
                  std::string& err) override {
        if (id >= max_scenes_) return (err = "too many scenes (" + std::to_string(max_scenes_) + ")", SORTGPU_ERR_ARG);
        SceneHost sh;
        int rc = prepare_scene(ctx_, acc, desc, true, sh.im);
        if (rc) return (err = sortgpu_last_error(ctx_), rc);
        if (sh.im.tree.valid) {
            if (width_ && sh.im.tree.width != width_)
                return (err = "scene node width " + std::to_string(sh.im.tree.width) + " differs from the server's " +
                              std::to_string(width_),
                        SORTGPU_ERR_ARG);
            width_ = sh.im.tree.width;
        }
        if (sh.im.tree.max_stack > trace_max_stack_supported())
            return (err = "tree too deep for the traversal stack", SORTGPU_ERR_ARG);
        max_stack_ = std::max(max_stack_, sh.im.tree.max_stack);
        sh.bytes = sh.im.block_bytes();
        *bytes = sh.bytes;
        scenes_.push_back(std::move(sh));
        return SORTGPU_OK;
    }

    int load_scene(uint32_t id, std::string& err) override {
        ctx_bind(ctx_);
        SceneHost& sh = scenes_[id];
        hipError_t e = hipMalloc(&sh.block, sh.bytes);
        if (e != hipSuccess) {
            sh.block = nullptr;
            return (err = hip_msg("hipMalloc(scene)", e), hip_rc(e));
        }
        SceneDev d;
        int rc = sh.im.place(ctx_, sh.block, d, s_up_);
        if (rc) {
            hipStreamSynchronize(s_up_);
            hipFree(sh.block);
            sh.block = nullptr;
            return (err = sortgpu_last_error(ctx_), rc);
        }
        h_scenes_[id] = d;
        h_trees_[id] = d.tree;
        h_transps_[id] = d.transp;
        if ((e = hipMemcpyAsync(d_scenes_ + id, h_scenes_ + id, sizeof(SceneDev), hipMemcpyHostToDevice, s_up_)) !=
                hipSuccess ||
            (e = hipMemcpyAsync(d_trees_ + id, h_trees_ + id, sizeof(TreeView), hipMemcpyHostToDevice, s_up_)) !=
                hipSuccess ||
            (e = hipMemcpyAsync(d_transps_ + id, h_transps_ + id, sizeof(const float*), hipMemcpyHostToDevice,
                                s_up_)) != hipSuccess ||
            (e = hipEventRecord(ev_up_, s_up_)) != hipSuccess)
            return (err = hip_msg("scene table update", e), hip_rc(e));
        return SORTGPU_OK;
    }

    void unload_scene(uint32_t id) override {
        ctx_bind(ctx_);
        SceneHost& sh = scenes_[id];
        if (!sh.block) return;
        sync_all();
        hipFree(sh.block);
        sh.block = nullptr;
    }

    int alloc_film(Request& r, std::string& err) override {
        ctx_bind(ctx_);
        void* p = nullptr;
        auto it = cache_.find(r.film_bytes);
        if (it != cache_.end()) {
            p = it->second;
            cached_ -= it->first;
            cache_.erase(it);
        } else {
            hipError_t e = hipMalloc(&p, r.film_bytes);
            if (e == hipErrorOutOfMemory && !cache_.empty()) {
                trim_cache();
                e = hipMalloc(&p, r.film_bytes);
            }
            if (e != hipSuccess) return (err = hip_msg("hipMalloc(film)", e), hip_rc(e));
        }
        r.film = p;  // cleared by the request's first wave
        return SORTGPU_OK;
    }

    void free_film(Request& r) override {
        if (!r.film) return;
        cache_.emplace(r.film_bytes, r.film);
        cached_ += r.film_bytes;
        r.film = nullptr;
    }

    size_t cached_bytes() const override { return cached_; }

    void trim_cache() override {
        if (cache_.empty()) return;
        ctx_bind(ctx_);
        sync_all();
        for (auto& kv : cache_) hipFree(kv.second);
        cache_.clear();
        cached_ = 0;
    }

    int launch(const Wave& w, std::string& err) override {
        ctx_bind(ctx_);
        WaveDesc wd;
        std::memset(&wd, 0, sizeof(wd));
        ChunkTable& ct = wd.ct;
        ReqDev* table = wd.reqs;
        if (w.chunks.size() > (size_t)kMaxChunks) return (err = "too many requests in a wave", SORTGPU_ERR_ARG);
        uint32_t path_base = 0, pix_base = 0;
        for (size_t k = 0; k < w.chunks.size(); k++) {
            const Chunk& c = w.chunks[k];
            const Request& r = *c.req;
            ReqDev& R = table[k];
            int rc = make_camera(ctx_, &r.p.camera, r.p.width, r.p.height, R.cam);
            if (rc) return (err = sortgpu_last_error(ctx_), rc);
            R.scene = r.p.scene;
            R.seed = fold_seed(r.p.seed);
            R.max_depth = r.p.max_depth;
            R.npix = r.npix;
            R.film = static_cast<float4*>(r.film);
            R.clamp = scenes_[r.p.scene].im.dev.clamp;
            ct.c[k] = ChunkDev{(uint32_t)k, c.sample0, c.nsamp, r.npix, path_base, pix_base};
            path_base += r.npix * c.nsamp;
            pix_base += r.npix;
        }
        if (path_base > cap_) return (err = "wave exceeds the workspace", SORTGPU_ERR_ARG);
        ct.n = (uint32_t)w.chunks.size();
        ct.total_paths = path_base;
        ct.total_pix = pix_base;
        const uint32_t k = (uint32_t)(w.seq % ns_);
        hipStream_t s = s_[k];
        Ev ev{w.seq, take_event(), take_event()};
        if (!ev.a || !ev.b) return (err = "hipEventCreate failed", SORTGPU_ERR_HIP);
        hipError_t e = hipEventRecord(ev.a, s);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, ev_up_, 0);
        for (size_t c = 0; c < w.chunks.size() && e == hipSuccess; c++)
            if (w.chunks[c].sample0 == 0) e = pt_zero_f4(table[c].film, table[c].npix, s);
        // the stream's own descriptor: its previous wave (same stream) has finished reading it
        const WaveDesc* d_wave = nullptr;
        int rc = SORTGPU_OK;
        if (e == hipSuccess) rc = ring_.push(ctx_, wd, s, d_desc_[k], &d_wave);
        if (e == hipSuccess && rc == SORTGPU_OK) {
            // a wave holding a scene with transparent materials answers its shadow rays with the
            // transparent-shadow walk (opaque scenes' rays in it get the same results as any-hit)
            bool att = false;
            for (const Chunk& c : w.chunks) att |= !scenes_[c.req->p.scene].im.transp.empty();
            PtLaunch L = wave_launch(k, w.max_depth, att);
            L.paths_cap = ct.total_paths;
            L.pix_cap = ct.total_pix;
            if (graphs_) {
                hipGraphExec_t x = nullptr;
                if ((rc = wave_graph(k, w.max_depth, att, &x)) == SORTGPU_OK) e = hipGraphLaunch(x, s);
            } else {
                e = pt_wave_paths(L, s);
            }
            // film updates in wave order: wait for the previous wave's
            if (e == hipSuccess && rc == SORTGPU_OK && film_recorded_) e = hipStreamWaitEvent(s, ev_film_, 0);
            if (e == hipSuccess && rc == SORTGPU_OK) e = pt_wave_film(L, s);
            if (e == hipSuccess && rc == SORTGPU_OK && (e = hipEventRecord(ev_film_, s)) == hipSuccess)
                film_recorded_ = true;
            if (e == hipSuccess && rc == SORTGPU_OK) rc = ring_.done(ctx_, s);
        }
        // the end event is recorded whatever happened, so the wave retires in order
        hipError_t e2 = hipEventRecord(ev.b, s);
        {
            std::lock_guard<std::mutex> l(ev_mu_);
            evs_.push_back(ev);
        }
        if (e != hipSuccess) return (err = hip_msg("wave launch", e), hip_rc(e));
        if (rc) return (err = sortgpu_last_error(ctx_), rc);
        if (e2 != hipSuccess) return (err = hip_msg("hipEventRecord", e2), hip_rc(e2));
        return SORTGPU_OK;
    }

    // The blocking form runs without the engine's lock while other threads may launch waves: the
    // event list is read and changed under ev_mu_ only, and the events of wave `seq` stay out of
    // the spare list (so no launch can record them again) until this call retires it.
    int wait(uint64_t seq, bool block, bool* done, std::string& err) override {
        ctx_bind(ctx_);
        *done = false;
        Ev ev;
        {
            std::lock_guard<std::mutex> l(ev_mu_);
            if (evs_.empty() || evs_.front().seq != seq)
                return (err = "wave bookkeeping out of order", SORTGPU_ERR_STATE);
            ev = evs_.front();
        }
        hipError_t e = block ? hipEventSynchronize(ev.b) : hipEventQuery(ev.b);
        if (e == hipErrorNotReady) return SORTGPU_OK;
        if (e != hipSuccess) return (err = hip_msg("wave", e), hip_rc(e));
        float ms = 0.0f;
        const bool timed = hipEventElapsedTime(&ms, ev.a, ev.b) == hipSuccess;
        {
            std::lock_guard<std::mutex> l(ev_mu_);
            if (evs_.empty() || evs_.front().seq != seq)
                return (err = "wave bookkeeping out of order", SORTGPU_ERR_STATE);
            if (timed) wave_us_ += (uint64_t)(ms * 1000.0f);
            evs_.pop_front();
            spare_.push_back(ev.a);
            spare_.push_back(ev.b);
        }
        uint32_t flags = 0;
        if ((e = hipMemcpy(&flags, d_error_, sizeof(flags), hipMemcpyDeviceToHost)) != hipSuccess)
            return (err = hip_msg("error flags", e), hip_rc(e));
        if (flags) return (err = "traversal error flags " + std::to_string(flags), SORTGPU_ERR_STATE);
        *done = true;
        return SORTGPU_OK;
    }

    int read_film(const Request& r, std::vector<float>& rgb, std::string& err) override {
        ctx_bind(ctx_);
        stage_.resize((size_t)r.npix * 4);
        hipError_t e = hipMemcpy(stage_.data(), r.film, r.film_bytes, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return (err = hip_msg("film copy", e), hip_rc(e));
        resolve(stage_.data(), r.npix, rgb);
        return SORTGPU_OK;
    }

    void counters(uint64_t out[3]) override {
        ctx_bind(ctx_);
        out[0] = out[1] = 0;
        for (uint32_t i = 0; i < ns_; i++) {
            unsigned long long t[2] = {0, 0};
            hipMemcpy(t, ws_[i].q.totals, sizeof(t), hipMemcpyDeviceToHost);
            out[0] += t[0];
            out[1] += t[1];
        }
        std::lock_guard<std::mutex> l(ev_mu_);
        out[2] = wave_us_;
    }

  private:
    struct SceneHost {
        SceneImage im;
        size_t bytes = 0;
        void* block = nullptr;
    };
    struct Ev {
        uint64_t seq;
        hipEvent_t a, b;
    };

    // Launch parameters of the waves of stream k: its workspace, its descriptor and its own work
    // counter (waves of one stream run one after another).
    PtLaunch wave_launch(uint32_t k, uint32_t depth, bool att) const {
        PtLaunch L;
        std::memset(&L, 0, sizeof(L));
        L.trace.counter = d_counters_ + k * kCounterStride;
        L.trace.error = d_error_;
        apply_tuning(ctx_, L.trace);
        L.scenes = d_scenes_;
        L.wave = d_desc_[k];
        L.trees = d_trees_;
        L.n_trees = max_scenes_;
        L.tree_width = width_ ? width_ : 8;
        L.tree_max_stack = max_stack_;
        L.ps = ws_[k].ps;
        L.q = ws_[k].q;
        L.max_depth = depth;
        L.num_cus = ctx_->num_cus;
        L.transps = att ? d_transps_ : nullptr;
        L.sort_min = shade_sort_min(ctx_->tuning);
        return L;
    }

    // The paths half of a wave on stream k captured as a graph (raygen + `depth` bounces, grids
    // sized for the workspace's capacity; the kernels read the wave's exact counts from its
    // descriptor). Captured on first use per depth, again when a scene changes the tree table's
    // width or stack depth or the context's tuning changes. The graph holds kernel launches only.
    int wave_graph(uint32_t k, uint32_t depth, bool att, hipGraphExec_t* out) {
        GraphKey key;
        std::memset(&key, 0, sizeof(key));
        key.tuning = ctx_->tuning;
        key.width = width_;
        key.max_stack = max_stack_;
        if (std::memcmp(&key, &graph_key_, sizeof(key)) != 0) {
            drop_graphs();
            graph_key_ = key;
        }
        const uint32_t gk = 2 * depth + (att ? 1u : 0u);
        auto it = gexec_[k].find(gk);
        if (it != gexec_[k].end()) {
            *out = it->second;
            return SORTGPU_OK;
        }
        hipStream_t s = s_[k];
        PtLaunch L = wave_launch(k, depth, att);
        L.paths_cap = cap_;
        L.pix_cap = cap_;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
        if (e != hipSuccess) return hip_fail(ctx_, e, "hipStreamBeginCapture(wave)");
        const hipError_t e1 = pt_wave_paths(L, s);
        e = hipStreamEndCapture(s, &g);
        if (e1 != hipSuccess || e != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return hip_fail(ctx_, e1 != hipSuccess ? e1 : e, "wave graph capture");
        }
        hipGraphExec_t x = nullptr;
        e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) return hip_fail(ctx_, e, "hipGraphInstantiate(wave)");
        gexec_[k][gk] = x;
        *out = x;
        return SORTGPU_OK;
    }

    void drop_graphs() {
        bool any = false;
        for (uint32_t i = 0; i < ns_; i++) any |= !gexec_[i].empty();
        if (!any) return;
        sync_all();
        for (uint32_t i = 0; i < ns_; i++) {
            for (auto& kv : gexec_[i]) hipGraphExecDestroy(kv.second);
            gexec_[i].clear();
        }
    }
    hipEvent_t take_event() {
        std::lock_guard<std::mutex> l(ev_mu_);
        if (!spare_.empty()) {
            hipEvent_t e = spare_.back();
            spare_.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        return hipEventCreate(&e) == hipSuccess ? e : nullptr;
    }
    void sync_all() {
        for (uint32_t i = 0; i < ns_; i++)
            if (s_[i]) hipStreamSynchronize(s_[i]);
        if (s_up_) hipStreamSynchronize(s_up_);
    }

    sortgpu_ctx* ctx_;
    uint32_t cap_, max_scenes_, ns_;
    bool graphs_;
    WaveDesc* d_desc_[kMaxStreams] = {};
    std::map<uint32_t, hipGraphExec_t> gexec_[kMaxStreams];  // key: 2 * depth + transparent-shadow form
    struct GraphKey {  // what a captured wave graph bakes in besides the depth
        sortgpu_tuning tuning;
        uint32_t width, max_stack;
    };
    GraphKey graph_key_{};
    hipStream_t s_[kMaxStreams] = {};
    hipStream_t s_up_ = nullptr;
    hipEvent_t ev_up_ = nullptr, ev_film_ = nullptr;
    bool film_recorded_ = false;
    PathWorkspace ws_[kMaxStreams];
    ReqRing ring_;
    SceneDev* d_scenes_ = nullptr;
    TreeView* d_trees_ = nullptr;
    SceneDev* h_scenes_ = nullptr;
    TreeView* h_trees_ = nullptr;
    const float** d_transps_ = nullptr;  // per scene: its transparency table on the device, or null
    const float** h_transps_ = nullptr;
    uint32_t* d_counters_ = nullptr;
    uint32_t* d_error_ = nullptr;
    uint32_t slot_ = 0;
    std::vector<SceneHost> scenes_;
    uint32_t width_ = 0, max_stack_ = 1;
    std::mutex ev_mu_;  // guards evs_, spare_ and wave_us_ (the blocking wait runs unlocked)
    std::deque<Ev> evs_;
    std::vector<hipEvent_t> spare_;
    std::multimap<size_t, void*> cache_;
    size_t cached_ = 0, fixed_ = 0;
    uint64_t wave_us_ = 0;
    std::vector<float> stage_;
};

// splitmix64 finaliser
inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

class HostDevice : public Device {
  public:
    HostDevice(sortgpu_ctx* ctx, uint32_t wave_paths, uint32_t polls) : ctx_(ctx), cap_(wave_paths), polls_(polls) {}
    ~HostDevice() override {
        for (auto* f : films_) delete f;
    }
    size_t fixed_bytes() const override { return PathWorkspace::bytes_for(cap_); }
    uint32_t wave_paths() const override { return cap_; }
    int add_scene(const sortgpu_accel* acc, const sortgpu_scene_desc* desc, uint32_t, size_t* bytes,
                  std::string& err) override {
        SceneImage im;
        int rc = prepare_scene(ctx_, acc, desc, true, im);
        if (rc) return (err = sortgpu_last_error(ctx_), rc);
        *bytes = im.block_bytes();
        return SORTGPU_OK;
    }
    int load_scene(uint32_t, std::string&) override { return SORTGPU_OK; }
    void unload_scene(uint32_t) override {}
    int alloc_film(Request& r, std::string&) override {
        auto* f = new std::vector<float>((size_t)r.npix * 4, 0.0f);
        films_.push_back(f);
        r.film = f;
        return SORTGPU_OK;
    }
    void free_film(Request& r) override {
        auto* f = static_cast<std::vector<float>*>(r.film);
        films_.erase(std::remove(films_.begin(), films_.end(), f), films_.end());
        delete f;
        r.film = nullptr;
    }
    size_t cached_bytes() const override { return 0; }
    void trim_cache() override {}
    int launch(const Wave& w, std::string&) override {
        for (const Chunk& c : w.chunks) {
            const Request& r = *c.req;
            float* acc = static_cast<std::vector<float>*>(r.film)->data();
            for (uint32_t s = c.sample0; s < c.sample0 + c.nsamp; s++)
                for (uint32_t pix = 0; pix < r.npix; pix++) {
                    const uint64_t h = mix64(r.p.seed ^ mix64(((uint64_t)r.p.scene << 48) ^ ((uint64_t)s << 24) ^ pix));
                    float* a = acc + 4 * (size_t)pix;
                    a[0] += (float)(h & 0xFFFF) * (1.0f / 65536.0f);
                    a[1] += (float)((h >> 16) & 0xFFFF) * (1.0f / 65536.0f);
                    a[2] += (float)((h >> 32) & 0xFFFF) * (1.0f / 65536.0f);
                    a[3] += 1.0f;
                }
        }
        std::lock_guard<std::mutex> l(mu_);
        ext_ += w.paths;
        pending_.push_back({w.seq, polls_});
        return SORTGPU_OK;
    }
    int wait(uint64_t seq, bool block, bool* done, std::string& err) override {
        std::lock_guard<std::mutex> l(mu_);
        *done = false;
        if (pending_.empty() || pending_.front().first != seq)
            return (err = "wave bookkeeping out of order", SORTGPU_ERR_STATE);
        if (!block && pending_.front().second > 0) {
            pending_.front().second--;
            return SORTGPU_OK;
        }
        pending_.pop_front();
        *done = true;
        return SORTGPU_OK;
    }
    int read_film(const Request& r, std::vector<float>& rgb, std::string&) override {
        resolve(static_cast<std::vector<float>*>(r.film)->data(), r.npix, rgb);
        return SORTGPU_OK;
    }
    void counters(uint64_t out[3]) override {
        std::lock_guard<std::mutex> l(mu_);
        out[0] = ext_;
        out[1] = 0;
        out[2] = 0;
    }

  private:
    sortgpu_ctx* ctx_;
    uint32_t cap_, polls_;
    std::vector<std::vector<float>*> films_;
    std::mutex mu_;  // guards pending_ and ext_ (the engine's blocking wait runs unlocked)
    std::deque<std::pair<uint64_t, uint32_t>> pending_;
    uint64_t ext_ = 0;
};

}  // namespace

int make_hip_device(sortgpu_ctx* ctx, uint32_t wave_paths, uint32_t max_scenes, uint32_t streams, bool graphs,
                    std::unique_ptr<Device>& out, std::string& err) {
    std::unique_ptr<HipDevice> d(new HipDevice(ctx, wave_paths, max_scenes, streams, graphs));
    int rc = d->init(err);
    if (rc) return rc;
    out = std::move(d);
    return SORTGPU_OK;
}

std::unique_ptr<Device> make_host_device(sortgpu_ctx* ctx, uint32_t wave_paths, uint32_t polls) {
    return std::unique_ptr<Device>(new HostDevice(ctx, wave_paths, polls));
}

}  // namespace serve
}  // namespace sortgpu
