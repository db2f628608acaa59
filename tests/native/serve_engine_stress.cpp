// Stress driver of the serving engine with a minimal in-process device, built by
// tests/test_serve_sanitizers.py under ThreadSanitizer and under AddressSanitizer + UBSan.
//
// Several client threads submit, wait, cancel, release and read results while the engine's worker
// (or, in manual mode, the clients' own wait calls) admits, plans, launches and retires waves; the device completes waves on a thread of its own
// after a short delay and fails a scripted fraction of film allocations. Every finished image is
// checked against the value the device's per-path hash gives in sample order, so a scheduling
// race that reorders or drops samples fails the run even when no sanitizer fires.
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "serve_engine.h"

using namespace sortgpu::serve;

namespace {

uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

float path_value(uint64_t seed, uint32_t scene, uint32_t pix, uint32_t s) {
    return (float)(mix64(seed ^ mix64(((uint64_t)scene << 48) ^ ((uint64_t)s << 24) ^ pix)) & 0xFFFF) * (1.0f / 65536.0f);
}

// Waves complete on a device thread in launch order, a little after launch.
class ThreadDevice : public Device {
  public:
    explicit ThreadDevice(uint32_t cap) : cap_(cap), th_([this] { run(); }) {}
    ~ThreadDevice() override {
        {
            std::lock_guard<std::mutex> l(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        th_.join();
        for (auto* f : films_) delete f;
    }
    size_t fixed_bytes() const override { return 1 << 20; }
    uint32_t wave_paths() const override { return cap_; }
    int add_scene(const sortgpu_accel*, const sortgpu_scene_desc*, uint32_t, size_t* bytes, std::string&) override {
        *bytes = 300000;
        return SORTGPU_OK;
    }
    int load_scene(uint32_t, std::string&) override { return SORTGPU_OK; }
    void unload_scene(uint32_t) override {}
    int alloc_film(Request& r, std::string& err) override {
        if (++allocs_ % 17 == 0) {
            err = "scripted allocation failure";
            return SORTGPU_ERR_OOM;
        }
        auto* f = new std::vector<float>((size_t)r.npix * 4, 0.0f);
        std::lock_guard<std::mutex> l(mu_);
        films_.push_back(f);
        r.film = f;
        return SORTGPU_OK;
    }
    void free_film(Request& r) override {
        auto* f = static_cast<std::vector<float>*>(r.film);
        std::lock_guard<std::mutex> l(mu_);
        for (auto& p : films_)
            if (p == f) {
                delete p;
                p = films_.back();
                films_.pop_back();
                break;
            }
        r.film = nullptr;
    }
    size_t cached_bytes() const override { return 0; }
    void trim_cache() override {}
    int launch(const Wave& w, std::string&) override {
        // copy what the device needs: the engine may retire and reuse request objects only after
        // this wave is reported done, so film pointers stay valid until then
        Job j;
        j.seq = w.seq;
        for (const Chunk& c : w.chunks)
            j.parts.push_back({static_cast<std::vector<float>*>(c.req->film), c.req->p.seed, c.req->p.scene,
                               c.req->npix, c.sample0, c.nsamp});
        std::lock_guard<std::mutex> l(mu_);
        jobs_.push_back(std::move(j));
        cv_.notify_all();
        return SORTGPU_OK;
    }
    int wait(uint64_t seq, bool block, bool* done, std::string&) override {
        std::unique_lock<std::mutex> l(mu_);
        if (block) cv_done_.wait(l, [&] { return done_seq_ >= seq; });
        *done = done_seq_ >= seq;
        return SORTGPU_OK;
    }
    int read_film(const Request& r, std::vector<float>& rgb, std::string&) override {
        const float* a = static_cast<std::vector<float>*>(r.film)->data();
        rgb.resize((size_t)r.npix * 3);
        for (uint32_t i = 0; i < r.npix; i++) {
            const float inv = a[4 * i + 3] > 0.0f ? 1.0f / a[4 * i + 3] : 0.0f;
            for (int c = 0; c < 3; c++) rgb[3 * i + c] = a[4 * i + c] * inv;
        }
        return SORTGPU_OK;
    }
    void counters(uint64_t out[3]) override { out[0] = out[1] = out[2] = 0; }

  private:
    struct Part {
        std::vector<float>* film;
        uint64_t seed;
        uint32_t scene, npix, s0, ns;
    };
    struct Job {
        uint64_t seq;
        std::vector<Part> parts;
    };
    void run() {
        std::unique_lock<std::mutex> l(mu_);
        for (;;) {
            cv_.wait(l, [&] { return stop_ || !jobs_.empty(); });
            if (jobs_.empty()) return;
            Job j = std::move(jobs_.front());
            jobs_.pop_front();
            l.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(200));
            for (const Part& p : j.parts)
                for (uint32_t s = p.s0; s < p.s0 + p.ns; s++)
                    for (uint32_t pix = 0; pix < p.npix; pix++) {
                        float* a = p.film->data() + 4 * (size_t)pix;
                        const float v = path_value(p.seed, p.scene, pix, s);
                        a[0] += v;
                        a[1] += v * 0.5f;
                        a[2] += v * 0.25f;
                        a[3] += 1.0f;
                    }
            l.lock();
            done_seq_ = j.seq;
            cv_done_.notify_all();
        }
    }
    uint32_t cap_;
    std::atomic<uint64_t> allocs_{0};
    std::mutex mu_;
    std::condition_variable cv_, cv_done_;
    std::deque<Job> jobs_;
    std::vector<std::vector<float>*> films_;
    uint64_t done_seq_ = 0;
    bool stop_ = false;
    std::thread th_;
};

bool check_image(const std::vector<float>& rgb, uint64_t seed, uint32_t scene, uint32_t npix, uint32_t spp) {
    for (uint32_t pix = 0; pix < npix; pix++) {
        float acc = 0.0f;
        for (uint32_t s = 0; s < spp; s++) acc += path_value(seed, scene, pix, s);
        const float inv = 1.0f / (float)spp;
        if (rgb[3 * pix] != acc * inv) return false;
    }
    return true;
}

}  // namespace

// One run of six clients against one engine. manual: no worker thread; the clients' own wait and
// drain calls drive the engine, so several threads poll it at once and take turns waiting on the
// device with the engine's lock released.
int run(bool manual) {
    Options o;
    o.budget = (1 << 20) + 4 * 300000 + 64 * 64 * 16 * 6;  // workspace + 4 scenes + six 64^2 films
    o.max_active = 6;
    o.inflight = 3;
    o.manual = manual;
    Engine eng(std::unique_ptr<Device>(new ThreadDevice(4096)), o);
    std::string err;
    if (eng.start(err)) {
        std::fprintf(stderr, "start: %s\n", err.c_str());
        return 2;
    }
    for (int i = 0; i < 6; i++) {
        uint32_t id;
        if (eng.add_scene(nullptr, nullptr, &id, err)) return 2;
    }
    std::atomic<int> bad{0}, done{0}, failed{0}, cancelled{0};
    auto client = [&](int t) {
        std::mt19937 g(1234 + t);
        for (int k = 0; k < 40; k++) {
            sortgpu_request r;
            std::memset(&r, 0, sizeof(r));
            r.scene = g() % 6;
            r.width = 8 + g() % 57;
            r.height = 8 + g() % 57;
            r.spp = 1 + g() % 12;
            r.max_depth = 8;
            r.seed = ((uint64_t)t << 32) | (uint64_t)k;
            r.camera.eye[2] = -1.0f;
            r.camera.up[1] = 1.0f;
            r.camera.fov_y = 0.7f;
            if (g() % 10 == 0) r.timeout_s = 1e-4;
            uint64_t id;
            std::string e;
            if (eng.submit(r, &id, e)) {
                bad++;
                continue;
            }
            if (g() % 7 == 0) eng.cancel(id);
            sortgpu_request_info info;
            eng.wait(id, 30.0, &info);
            if (info.state == SORTGPU_REQ_DONE) {
                std::vector<float> rgb((size_t)r.width * r.height * 3);
                if (eng.result(id, rgb.data(), rgb.size(), e) ||
                    !check_image(rgb, r.seed, r.scene, r.width * r.height, r.spp))
                    bad++;
                done++;
            } else if (info.state == SORTGPU_REQ_FAILED) {
                failed++;
            } else if (info.state == SORTGPU_REQ_CANCELLED || info.state == SORTGPU_REQ_TIMED_OUT) {
                cancelled++;
            } else {
                bad++;  // not finished within the wait
            }
            eng.release(id);
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < 6; t++) th.emplace_back(client, t);
    for (auto& x : th) x.join();
    if (eng.drain(30.0)) bad++;
    uint64_t st[16];
    eng.stats(st);
    std::printf("%s: done %d failed %d cancelled %d bad %d waves %llu evictions %llu peak %llu budget %llu\n",
                manual ? "manual" : "worker", done.load(), failed.load(), cancelled.load(), bad.load(),
                (unsigned long long)st[0], (unsigned long long)st[5], (unsigned long long)st[8],
                (unsigned long long)st[14]);
    if (st[8] > st[14]) bad++;
    return bad.load() == 0 && done.load() > 100 ? 0 : 1;
}

int main() {
    const int a = run(false);
    const int b = run(true);
    return a | b;
}
