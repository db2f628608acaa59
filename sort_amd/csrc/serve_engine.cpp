// Serving engine: admission, eviction, wave planning, completion, cancellation and timeouts.
#include "serve_engine.h"

#include <algorithm>
#include <cmath>

namespace sortgpu {
namespace serve {

namespace {
bool terminal(int s) { return s >= SORTGPU_REQ_DONE; }

std::chrono::steady_clock::time_point deadline_after(double timeout_s) {
    const double t = timeout_s < 0 ? 1e7 : std::min(timeout_s, 1e7);
    return std::chrono::steady_clock::now() +
           std::chrono::duration_cast<std::chrono::steady_clock::duration>(std::chrono::duration<double>(t));
}
}  // namespace

Engine::Engine(std::unique_ptr<Device> dev, const Options& o) : dev_(std::move(dev)), o_(o) {
    t0_ = std::chrono::steady_clock::now();
    o_.max_active = std::max(1u, std::min(o_.max_active, 32u));
    o_.inflight = std::max(1u, std::min(o_.inflight, 4u));
}

Engine::~Engine() {
    {
        std::unique_lock<std::mutex> lk(mu_);
        stop_ = true;
        for (Request* r : std::vector<Request*>(queue_.begin(), queue_.end()))
            finish(r, SORTGPU_REQ_CANCELLED, SORTGPU_OK, "server shut down");
        for (Request* r : std::vector<Request*>(active_)) finish(r, SORTGPU_REQ_CANCELLED, SORTGPU_OK, "server shut down");
    }
    cv_work_.notify_all();
    if (worker_.joinable()) worker_.join();
    std::unique_lock<std::mutex> lk(mu_);
    while (!inflight_.empty()) poll(lk, true);
    reap();
    for (uint32_t i = 0; i < scenes_.size(); i++)
        if (scenes_[i].resident) dev_->unload_scene(i);
    cv_done_.notify_all();
}

double Engine::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

int Engine::start(std::string& err) {
    std::lock_guard<std::mutex> lk(mu_);
    used_ = dev_->fixed_bytes();
    peak_ = used_;
    if (used_ > o_.budget) {
        err = "the wave workspace (" + std::to_string(used_) + " bytes) exceeds the HBM budget (" +
              std::to_string(o_.budget) + ")";
        return SORTGPU_ERR_BUDGET;
    }
    started_ = true;
    if (!o_.manual) worker_ = std::thread([this] { loop(); });
    return SORTGPU_OK;
}

int Engine::add_scene(const sortgpu_accel* acc, const sortgpu_scene_desc* desc, uint32_t* id, std::string& err) {
    // registration is host-side work on the device object, serialised with the worker
    std::unique_lock<std::mutex> lk(mu_);
    const uint32_t sid = (uint32_t)scenes_.size();
    size_t bytes = 0;
    int rc = dev_->add_scene(acc, desc, sid, &bytes, err);
    if (rc) return rc;
    SceneState s;
    s.bytes = bytes;
    scenes_.push_back(s);
    *id = sid;
    return SORTGPU_OK;
}

int Engine::submit(const sortgpu_request& p, uint64_t* id, std::string& err) {
    std::unique_lock<std::mutex> lk(mu_);
    if (p.scene >= scenes_.size()) {
        err = "unknown scene " + std::to_string(p.scene);
        return SORTGPU_ERR_ARG;
    }
    if (p.width == 0 || p.height == 0 || p.spp == 0) {
        err = "width, height and spp must be positive";
        return SORTGPU_ERR_ARG;
    }
    const uint64_t npix = (uint64_t)p.width * p.height;
    if (npix > dev_->wave_paths()) {
        err = "image of " + std::to_string(npix) + " pixels exceeds the wave size (" +
              std::to_string(dev_->wave_paths()) + " paths)";
        return SORTGPU_ERR_ARG;
    }
    if (p.spp > (1u << 24) || p.max_depth == 0 || p.max_depth > 64) {
        err = "spp above 2^24, or max_depth outside 1..64";
        return SORTGPU_ERR_ARG;
    }
    const sortgpu_camera& c = p.camera;
    const float fx = c.target[0] - c.eye[0], fy = c.target[1] - c.eye[1], fz = c.target[2] - c.eye[2];
    const float ux = c.up[1] * fz - c.up[2] * fy, uy = c.up[2] * fx - c.up[0] * fz, uz = c.up[0] * fy - c.up[1] * fx;
    if (!(fx * fx + fy * fy + fz * fz > 0.0f) || !(ux * ux + uy * uy + uz * uz > 0.0f) || !(c.fov_y > 0.0f) ||
        !(c.fov_y < 3.14159f)) {
        err = "degenerate camera (eye == target, up parallel to the view, or fov outside (0, pi))";
        return SORTGPU_ERR_ARG;
    }
    if (broken_) {
        err = "server failed: " + broken_msg_;
        return SORTGPU_ERR_STATE;
    }
    std::unique_ptr<Request> r(new Request());
    r->id = next_id_++;
    r->p = p;
    r->npix = (uint32_t)npix;
    r->film_bytes = npix * 16;
    r->t_submit = now();
    r->deadline = p.timeout_s > 0 ? r->t_submit + p.timeout_s : 0.0;
    queue_.push_back(r.get());
    *id = r->id;
    reqs_[r->id] = std::move(r);
    lk.unlock();
    cv_work_.notify_all();
    return SORTGPU_OK;
}

Request* Engine::find(uint64_t id) {
    auto it = reqs_.find(id);
    return it == reqs_.end() || it->second->released ? nullptr : it->second.get();
}

void Engine::drop_film(Request* r) {
    if (!r->film) return;
    dev_->free_film(*r);
    r->film = nullptr;
    used_ -= r->film_bytes;
}

// Moves a request to a terminal state. Its film is freed at once unless a wave in flight still
// writes it; then it waits among the zombies until that wave is retired.
void Engine::finish(Request* r, int state, int status, const std::string& msg) {
    const bool was_active = r->state == SORTGPU_REQ_ACTIVE;
    r->state = state;
    r->status = status;
    if (!msg.empty()) r->msg = msg;
    r->t_done = now();
    if (was_active) {
        active_.erase(std::remove(active_.begin(), active_.end(), r), active_.end());
        SceneState& s = scenes_[r->p.scene];
        if (s.refs) s.refs--;
    } else {
        queue_.erase(std::remove(queue_.begin(), queue_.end(), r), queue_.end());
    }
    if (r->film) {
        if (r->last_wave > done_seq_)
            zombies_.push_back(r);
        else
            drop_film(r);
    }
    switch (state) {
        case SORTGPU_REQ_DONE: n_done_++; break;
        case SORTGPU_REQ_FAILED: n_failed_++; break;
        case SORTGPU_REQ_CANCELLED: n_cancel_++; break;
        case SORTGPU_REQ_TIMED_OUT: n_timeout_++; break;
    }
    cv_done_.notify_all();
}

// Frees zombie films whose waves are done and forgets released requests nothing refers to.
void Engine::reap() {
    std::vector<Request*> keep;
    for (Request* r : zombies_) {
        if (r->last_wave <= done_seq_)
            drop_film(r);
        else
            keep.push_back(r);
    }
    zombies_.swap(keep);
    for (auto it = reqs_.begin(); it != reqs_.end();) {
        Request* r = it->second.get();
        if (r->released && terminal(r->state) && !r->film && r->last_wave <= done_seq_)
            it = reqs_.erase(it);
        else
            ++it;
    }
}

int Engine::cancel(uint64_t id) {
    std::unique_lock<std::mutex> lk(mu_);
    Request* r = find(id);
    if (!r) return SORTGPU_ERR_ARG;
    if (!terminal(r->state)) finish(r, SORTGPU_REQ_CANCELLED, SORTGPU_OK, "cancelled");
    lk.unlock();
    cv_work_.notify_all();
    return SORTGPU_OK;
}

static void fill_info(const Request* r, sortgpu_request_info* out) {
    out->state = r->state;
    out->status = r->status;
    out->samples_done = r->done;
    out->spp = r->p.spp;
    out->t_submit = r->t_submit;
    out->t_admit = r->t_admit;
    out->t_done = terminal(r->state) ? r->t_done : 0.0;
}

int Engine::info(uint64_t id, sortgpu_request_info* out) {
    std::lock_guard<std::mutex> lk(mu_);
    Request* r = find(id);
    if (!r) return SORTGPU_ERR_ARG;
    fill_info(r, out);
    return SORTGPU_OK;
}

int Engine::wait(uint64_t id, double timeout_s, sortgpu_request_info* out) {
    std::unique_lock<std::mutex> lk(mu_);
    const auto t_end = deadline_after(timeout_s);
    for (;;) {
        Request* r = find(id);
        if (!r) return SORTGPU_ERR_ARG;
        if (terminal(r->state) || std::chrono::steady_clock::now() >= t_end) {
            if (out) fill_info(r, out);
            return SORTGPU_OK;
        }
        if (o_.manual) {
            std::string err;
            iterate(lk, true);
        } else {
            cv_done_.wait_until(lk, t_end);
        }
    }
}

int Engine::result(uint64_t id, float* rgb, uint64_t n_floats, std::string& err) {
    std::lock_guard<std::mutex> lk(mu_);
    Request* r = find(id);
    if (!r) {
        err = "unknown request";
        return SORTGPU_ERR_ARG;
    }
    if (r->state != SORTGPU_REQ_DONE) {
        err = "request is not done";
        return SORTGPU_ERR_STATE;
    }
    if (n_floats < r->rgb.size() || !rgb) {
        err = "result buffer holds " + std::to_string(n_floats) + " floats, " + std::to_string(r->rgb.size()) +
              " needed";
        return SORTGPU_ERR_ARG;
    }
    std::copy(r->rgb.begin(), r->rgb.end(), rgb);
    return SORTGPU_OK;
}

int Engine::message(uint64_t id, std::string& msg) {
    std::lock_guard<std::mutex> lk(mu_);
    Request* r = find(id);
    if (!r) return SORTGPU_ERR_ARG;
    msg = r->msg;
    return SORTGPU_OK;
}

int Engine::release(uint64_t id) {
    std::unique_lock<std::mutex> lk(mu_);
    Request* r = find(id);
    if (!r) return SORTGPU_ERR_ARG;
    if (!terminal(r->state)) finish(r, SORTGPU_REQ_CANCELLED, SORTGPU_OK, "released");
    r->released = true;
    r->rgb.clear();
    r->rgb.shrink_to_fit();
    reap();
    lk.unlock();
    cv_work_.notify_all();
    return SORTGPU_OK;
}

void Engine::expire() {
    const double t = now();
    std::vector<Request*> late;
    for (Request* r : queue_)
        if (r->deadline > 0 && t >= r->deadline) late.push_back(r);
    for (Request* r : active_)
        if (r->deadline > 0 && t >= r->deadline) late.push_back(r);
    for (Request* r : late) finish(r, SORTGPU_REQ_TIMED_OUT, SORTGPU_OK, "timed out");
}

int Engine::reserve(size_t bytes, std::string& err) {
    alloc_count_++;
    if (o_.fail_alloc_at && alloc_count_ == o_.fail_alloc_at) {
        err = "injected allocation failure";
        return SORTGPU_ERR_OOM;
    }
    if (used_ + dev_->cached_bytes() + bytes > o_.budget) dev_->trim_cache();
    if (used_ + bytes > o_.budget) {
        err = "over the HBM budget";
        return SORTGPU_ERR_BUDGET;
    }
    return SORTGPU_OK;
}

// Evicts least-recently-used resident scenes that no active request uses and no wave in flight
// reads, until `need` more bytes fit. Returns whether they fit.
bool Engine::evict_for(size_t need, uint32_t keep_scene) {
    if (used_ + dev_->cached_bytes() + need > o_.budget) dev_->trim_cache();
    while (used_ + need > o_.budget) {
        int victim = -1;
        for (uint32_t i = 0; i < scenes_.size(); i++) {
            const SceneState& s = scenes_[i];
            if (!s.resident || s.refs || i == keep_scene || s.last_wave > done_seq_) continue;
            if (victim < 0 || s.lru < scenes_[victim].lru) victim = (int)i;
        }
        if (victim < 0) return false;
        dev_->unload_scene((uint32_t)victim);
        scenes_[victim].resident = false;
        used_ -= scenes_[victim].bytes;
        n_evict_++;
    }
    return true;
}

// Admits queued requests in arrival order while there is room: a free active slot, the scene
// resident (or loadable) and the film within budget. The head of the queue waits for memory held
// by running requests; a request that cannot fit even in an idle server fails.
void Engine::admit() {
    while (!queue_.empty() && active_.size() < o_.max_active) {
        Request* r = queue_.front();
        SceneState& s = scenes_[r->p.scene];
        const size_t need = (s.resident ? 0 : s.bytes) + r->film_bytes;
        if (!evict_for(need, r->p.scene)) {
            if (active_.empty() && inflight_.empty()) {
                finish(r, SORTGPU_REQ_FAILED, SORTGPU_ERR_BUDGET,
                       "needs " + std::to_string(need) + " bytes; the HBM budget is " + std::to_string(o_.budget) +
                           " with " + std::to_string(used_) + " in use");
                continue;
            }
            return;
        }
        std::string err;
        int rc;
        if (!s.resident) {
            if ((rc = reserve(s.bytes, err)) || (rc = dev_->load_scene(r->p.scene, err))) {
                finish(r, SORTGPU_REQ_FAILED, rc, "scene load failed: " + err);
                continue;
            }
            s.resident = true;
            used_ += s.bytes;
            n_loads_++;
        }
        if ((rc = reserve(r->film_bytes, err)) || (rc = dev_->alloc_film(*r, err))) {
            finish(r, SORTGPU_REQ_FAILED, rc, "film allocation failed: " + err);
            continue;
        }
        used_ += r->film_bytes;
        peak_ = std::max(peak_, used_);
        queue_.pop_front();
        r->state = SORTGPU_REQ_ACTIVE;
        r->t_admit = now();
        s.refs++;
        s.lru = ++tick_;
        active_.push_back(r);
    }
}

// Packs the next wave: one more sample per active request per round, oldest request first, while
// the wave has room; at most 32 requests per wave.
bool Engine::plan_and_launch() {
    std::vector<Request*> cand;
    for (Request* r : active_)
        if (r->planned < r->p.spp && cand.size() < 32) cand.push_back(r);
    if (cand.empty()) return false;
    const uint64_t cap = dev_->wave_paths();
    std::vector<uint32_t> take(cand.size(), 0);
    uint64_t paths = 0;
    for (bool progress = true; progress;) {
        progress = false;
        for (size_t i = 0; i < cand.size(); i++) {
            Request* r = cand[i];
            if (r->planned + take[i] < r->p.spp && paths + r->npix <= cap) {
                take[i]++;
                paths += r->npix;
                progress = true;
            }
        }
    }
    Wave w;
    w.seq = wave_seq_ + 1;
    for (size_t i = 0; i < cand.size(); i++)
        if (take[i]) {
            w.chunks.push_back(Chunk{cand[i], cand[i]->planned, take[i]});
            w.max_depth = std::max(w.max_depth, cand[i]->p.max_depth);
        }
    if (w.chunks.empty()) return false;
    w.paths = paths;
    std::string err;
    const int rc = dev_->launch(w, err);
    wave_seq_ = w.seq;
    for (const Chunk& c : w.chunks) {
        c.req->planned += c.nsamp;
        c.req->last_wave = w.seq;
        scenes_[c.req->p.scene].last_wave = w.seq;
        scenes_[c.req->p.scene].lru = ++tick_;
    }
    inflight_.push_back(std::move(w));
    n_waves_++;
    n_paths_ += paths;
    if (rc) {
        for (const Chunk& c : inflight_.back().chunks)
            if (c.req->state == SORTGPU_REQ_ACTIVE) finish(c.req, SORTGPU_REQ_FAILED, rc, "wave launch failed: " + err);
    }
    return true;
}

void Engine::retire(Wave& w) {
    for (const Chunk& c : w.chunks) {
        Request* r = c.req;
        if (r->state != SORTGPU_REQ_ACTIVE) continue;
        r->done += c.nsamp;
        if (r->done == r->p.spp) {
            std::string err;
            const int rc = dev_->read_film(*r, r->rgb, err);
            if (rc)
                finish(r, SORTGPU_REQ_FAILED, rc, "film read failed: " + err);
            else
                finish(r, SORTGPU_REQ_DONE, SORTGPU_OK, "");
        }
    }
}

// Retires finished waves in order; with block, waits for the oldest one (mu released meanwhile).
// One thread at a time may wait on the device: while it does (polling_), every other poll returns
// at once, or with block waits until that thread has retired what finished, so the device's wait
// is never entered twice and no wave is retired by two threads.
void Engine::poll(std::unique_lock<std::mutex>& lk, bool block) {
    if (polling_) {
        if (block) cv_done_.wait(lk, [this] { return !polling_; });
        return;
    }
    bool first = true;
    while (!inflight_.empty()) {
        const uint64_t seq = inflight_.front().seq;
        bool done = false;
        std::string err;
        int rc;
        if (block && first) {
            polling_ = true;
            lk.unlock();
            rc = dev_->wait(seq, true, &done, err);
            lk.lock();
            polling_ = false;
            cv_done_.notify_all();
            // nothing else retires waves while polling_ is set; the check keeps that an invariant
            if (!rc && (inflight_.empty() || inflight_.front().seq != seq)) break;
        } else {
            rc = dev_->wait(seq, false, &done, err);
        }
        first = false;
        if (rc) {
            // the device is unusable: fail everything in flight and refuse new work
            broken_ = true;
            broken_msg_ = err;
            for (Wave& w : inflight_)
                for (const Chunk& c : w.chunks)
                    if (!terminal(c.req->state)) finish(c.req, SORTGPU_REQ_FAILED, rc, "device failure: " + err);
            for (Request* r : std::vector<Request*>(active_)) finish(r, SORTGPU_REQ_FAILED, rc, "device failure: " + err);
            for (Request* r : std::vector<Request*>(queue_.begin(), queue_.end()))
                finish(r, SORTGPU_REQ_FAILED, rc, "device failure: " + err);
            done_seq_ = wave_seq_;
            inflight_.clear();
            return;
        }
        if (!done) break;
        Wave w = std::move(inflight_.front());
        inflight_.pop_front();
        done_seq_ = w.seq;
        retire(w);
    }
    reap();
}

bool Engine::iterate(std::unique_lock<std::mutex>& lk, bool block) {
    poll(lk, false);
    expire();
    if (!broken_) admit();
    while (!broken_ && inflight_.size() < o_.inflight && plan_and_launch()) {
    }
    if (block && !inflight_.empty()) {
        poll(lk, true);
        expire();
        if (!broken_) admit();
        while (!broken_ && inflight_.size() < o_.inflight && plan_and_launch()) {
        }
    }
    return !queue_.empty() || !active_.empty() || !inflight_.empty();
}

void Engine::loop() {
    std::unique_lock<std::mutex> lk(mu_);
    while (!stop_) {
        const bool busy = iterate(lk, false);
        if (stop_) break;
        if (inflight_.size() >= o_.inflight) {
            poll(lk, true);  // the window is full: nothing can launch before the oldest wave ends
        } else if (!inflight_.empty()) {
            // room in the window: launch new work as soon as it arrives, else poll again shortly
            cv_work_.wait_for(lk, std::chrono::microseconds(200));
        } else {
            // idle, or waiting for memory / a deadline
            cv_work_.wait_for(lk, std::chrono::milliseconds(busy ? 1 : 20));
        }
    }
}

int Engine::step(int* busy, std::string& err) {
    if (!o_.manual) {
        err = "sortgpu_server_step needs a server created in manual mode";
        return SORTGPU_ERR_STATE;
    }
    std::unique_lock<std::mutex> lk(mu_);
    const bool b = iterate(lk, true);
    if (busy) *busy = b ? 1 : 0;
    return SORTGPU_OK;
}

int Engine::drain(double timeout_s) {
    std::unique_lock<std::mutex> lk(mu_);
    const auto t_end = deadline_after(timeout_s);
    for (;;) {
        if (queue_.empty() && active_.empty() && inflight_.empty()) return SORTGPU_OK;
        if (std::chrono::steady_clock::now() >= t_end) return SORTGPU_ERR_STATE;
        if (o_.manual)
            iterate(lk, true);
        else
            cv_done_.wait_until(lk, std::min(t_end, std::chrono::steady_clock::now() + std::chrono::milliseconds(5)));
    }
}

void Engine::stats(uint64_t out[16]) {
    std::lock_guard<std::mutex> lk(mu_);
    uint64_t c[3] = {0, 0, 0};
    dev_->counters(c);
    uint32_t resident = 0;
    for (const SceneState& s : scenes_) resident += s.resident ? 1 : 0;
    out[0] = n_waves_;
    out[1] = n_paths_;
    out[2] = c[0];
    out[3] = c[1];
    out[4] = n_loads_;
    out[5] = n_evict_;
    out[6] = resident;
    out[7] = used_;
    out[8] = peak_;
    out[9] = n_done_;
    out[10] = n_failed_;
    out[11] = n_cancel_;
    out[12] = n_timeout_;
    out[13] = c[2];
    out[14] = o_.budget;
    out[15] = dev_->fixed_bytes();
}

}  // namespace serve
}  // namespace sortgpu
