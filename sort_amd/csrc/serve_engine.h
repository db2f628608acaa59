// Serving engine: a continuous-batching request loop over resident scenes.
//
// Requests (scene, camera, resolution, spp, depth, seed) arrive at any time. The engine admits
// them in arrival order while the HBM budget allows (loading the scene when it is not resident,
// evicting least-recently-used idle scenes to make room), then packs whole samples of every
// active request into waves of a fixed number of path slots: one sample per request per round,
// oldest first, until the wave is full. Each wave is one raygen + bounce loop + film update on
// the device; waves run in order on one stream, a few enqueued ahead of the host. A request's
// film is a sum of its samples in sample order, and every path's radiance depends only on
// (seed, pixel, sample), so a result does not depend on what else was in its waves.
//
// The device side sits behind `Device` so the same engine runs on the GPU or on a host-only
// stand-in (tests of admission, eviction, cancellation, timeouts and injected failures).
#pragma once
#include <stdint.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sortgpu.h"

namespace sortgpu {
namespace serve {

struct Request {
    uint64_t id = 0;
    sortgpu_request p{};
    uint32_t npix = 0;
    int state = SORTGPU_REQ_QUEUED;
    int status = SORTGPU_OK;
    std::string msg;
    uint32_t planned = 0, done = 0;  // samples put in waves / finished
    uint64_t last_wave = 0;          // last wave holding samples of this request (0 = none)
    void* film = nullptr;            // device film handle (null = none held)
    size_t film_bytes = 0;
    bool released = false;
    double t_submit = 0, t_admit = 0, t_done = 0, deadline = 0;
    std::vector<float> rgb;          // result (DONE)
};

struct Chunk {
    Request* req;
    uint32_t sample0, nsamp;
};

struct Wave {
    uint64_t seq = 0;
    uint64_t paths = 0;
    uint32_t max_depth = 0;
    std::vector<Chunk> chunks;
};

// Device side of the engine. The engine calls it with its lock held, except for one blocking
// wait at a time, which runs unlocked while other threads may call the rest.
class Device {
  public:
    virtual ~Device() {}
    virtual size_t fixed_bytes() const = 0;   // workspace etc., held for the server's life
    virtual uint32_t wave_paths() const = 0;  // path slots per wave
    // Host-side registration; returns the HBM bytes the scene takes when resident.
    virtual int add_scene(const sortgpu_accel* acc, const sortgpu_scene_desc* desc, uint32_t id, size_t* bytes,
                          std::string& err) = 0;
    virtual int load_scene(uint32_t id, std::string& err) = 0;
    virtual void unload_scene(uint32_t id) = 0;
    virtual int alloc_film(Request& r, std::string& err) = 0;  // zeroed; sets r.film
    virtual void free_film(Request& r) = 0;
    virtual size_t cached_bytes() const = 0;  // freed films kept for reuse
    virtual void trim_cache() = 0;
    virtual int launch(const Wave& w, std::string& err) = 0;  // asynchronous
    // Waves finish in launch order. Sets *done; with block, waits for wave `seq`.
    virtual int wait(uint64_t seq, bool block, bool* done, std::string& err) = 0;
    virtual int read_film(const Request& r, std::vector<float>& rgb, std::string& err) = 0;
    // [0] extension rays, [1] shadow rays, [2] device busy microseconds
    virtual void counters(uint64_t out[3]) = 0;
};

struct Options {
    uint64_t budget = 96ull << 30;
    uint32_t max_active = 32;
    uint32_t inflight = 2;
    bool manual = false;
    uint32_t fail_alloc_at = 0;
};

class Engine {
  public:
    Engine(std::unique_ptr<Device> dev, const Options& o);
    ~Engine();
    int start(std::string& err);  // checks the budget against fixed bytes; starts the worker
    int add_scene(const sortgpu_accel* acc, const sortgpu_scene_desc* desc, uint32_t* id, std::string& err);
    int submit(const sortgpu_request& r, uint64_t* id, std::string& err);
    int cancel(uint64_t id);
    int info(uint64_t id, sortgpu_request_info* out);
    int wait(uint64_t id, double timeout_s, sortgpu_request_info* out);
    int result(uint64_t id, float* rgb, uint64_t n_floats, std::string& err);
    int message(uint64_t id, std::string& msg);
    int release(uint64_t id);
    int step(int* busy, std::string& err);  // manual mode only
    int drain(double timeout_s);
    void stats(uint64_t out[16]);

  private:
    struct SceneState {
        size_t bytes = 0;
        bool resident = false;
        uint32_t refs = 0;       // active requests on it
        uint64_t lru = 0;
        uint64_t last_wave = 0;  // last wave that read it
    };
    double now() const;
    void loop();
    // One scheduling pass with mu held; may block (mu released) on the oldest wave.
    bool iterate(std::unique_lock<std::mutex>& lk, bool block);
    void poll(std::unique_lock<std::mutex>& lk, bool block);
    void retire(Wave& w);
    void finish(Request* r, int state, int status, const std::string& msg);
    void drop_film(Request* r);
    void reap();
    void expire();
    void admit();
    bool plan_and_launch();
    bool evict_for(size_t need, uint32_t keep_scene);
    int reserve(size_t bytes, std::string& err);
    Request* find(uint64_t id);

    std::unique_ptr<Device> dev_;
    Options o_;
    std::chrono::steady_clock::time_point t0_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::thread worker_;
    bool stop_ = false, started_ = false, broken_ = false;
    bool polling_ = false;  // a thread is blocked in dev_->wait with mu_ released
    std::string broken_msg_;
    std::map<uint64_t, std::unique_ptr<Request>> reqs_;
    std::deque<Request*> queue_;    // QUEUED, arrival order
    std::vector<Request*> active_;  // ACTIVE, admission order
    std::vector<Request*> zombies_; // finished requests whose film a wave in flight may still write
    std::deque<Wave> inflight_;
    std::vector<SceneState> scenes_;
    uint64_t next_id_ = 1, wave_seq_ = 0, done_seq_ = 0, tick_ = 0, alloc_count_ = 0;
    size_t used_ = 0, peak_ = 0;
    uint64_t n_waves_ = 0, n_paths_ = 0, n_loads_ = 0, n_evict_ = 0;
    uint64_t n_done_ = 0, n_failed_ = 0, n_cancel_ = 0, n_timeout_ = 0;
};

}  // namespace serve
}  // namespace sortgpu
