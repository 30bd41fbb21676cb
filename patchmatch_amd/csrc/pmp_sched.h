// pmp_sched.h -- the served form's scheduler, independent of the device.
//
// A scheduler thread loops over steps.  A step admits queued requests into
// free slots (keyframe maps charged against the memory budget), then hands
// the executor one stage of every slot in flight, each at its own stage:
//   stage 0                     random init + initial score (+ landmark seeding)
//   stage 1 .. last-1           red-black pass stage-1
//   stage last = 2 iter + 1     copy-back, gate, PostProcess, maps to the host
// so a request that arrives while others run joins at the next step.  Steps
// are issued two deep (executor descriptor sets 0 and 1 alternate); the
// completions of step k are collected after step k+1 is issued.
//
// Images live in a pool keyed by the caller's ids, with an LRU clock and a
// pin count (requests queued or in flight that read them); unpinned images
// are evicted oldest first when uploads or slot maps need room in the budget.
//
// Ending a request (done, failed, cancelled) waits until no issued step
// that includes it is still running: only then is its slot reused, its pins
// dropped and its host buffers handed back, and only then does it become
// visible to wait/fetch.  The reference's controls are the densify thread's
// request_reset / request_pause / request_terminate
// (depthmap_computation_module.cc:503-615); here reset cancels every request
// and empties the pool, and cancel ends one request.
//
// The device side is an Executor (pmp_server.hip on the GPU; a simulated one
// drives the scheduler in the CPU sanitizer tests).
#pragma once
#include <stdint.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pmp/pmp.h"

namespace pmp {
namespace sched {

// One request as the executor sees it.  out: host buffer of W*H*28 bytes
// (plane 16 B | depth | score | nghbr per pixel, each map contiguous), written
// by the step that finalizes the request; lm: n_lm landmark positions;
// prev_*: the warm start's previous result.
struct Job {
    uint64_t ticket = 0;
    pmp_keyframe kf{};
    pmp_params p{};
    int W = 0, H = 0;
    int last_stage = 1;
    const double* lm = nullptr;
    int64_t n_lm = 0;
    // warm start: the previous keyframe's plane (4 floats per pixel) and depth
    // maps of W x H pixels, in pinned host memory, and its pose (R row-major,
    // then t); prev_plane4 null: none
    const float* prev_plane4 = nullptr;
    const float* prev_depth = nullptr;
    double prev_pose[12] = {};
    uint8_t* out = nullptr;
};

struct StepItem {
    int slot;
    int stage;
    const Job* job;
};

class Executor {
  public:
    virtual ~Executor() = default;
    // Called once on the scheduler thread before its first step.
    virtual void bind_thread() {}
    // PMP_OK or PMP_EINVAL with a reason.
    virtual int check_params(const pmp_params& p, std::string& err) = 0;
    // Device bytes of a pooled W x H image (with its packed neighbour copy)
    // and of the keyframe maps of a W x H slot.
    virtual int64_t image_bytes(int W, int H, bool packed) const = 0;
    virtual int64_t map_bytes(int64_t pixels) const = 0;
    // Pool entries; each call has completed when it returns.
    virtual int upload(uint64_t id, const uint8_t* pixels, int W, int H, std::string& err) = 0;
    virtual int pack(uint64_t id, int wrap, std::string& err) = 0;
    virtual void drop(uint64_t id) = 0;
    // Keyframe maps of slot k for `pixels` pixels (replacing smaller ones).
    virtual int slot_grow(int k, int64_t pixels, std::string& err) = 0;
    // Pinned host buffers (nullptr on failure).
    virtual uint8_t* host_alloc(size_t bytes) = 0;
    virtual void host_free(uint8_t* p) = 0;
    // Issues one step into descriptor set w (0 or 1), asynchronously.
    // item_rc[k] != 0 leaves item k out (its request fails with that status
    // and item_err[k]); a non-zero return means the step could not be issued
    // at all (err).
    virtual int step_issue(int w, const std::vector<StepItem>& items, std::vector<int>& item_rc,
                           std::vector<std::string>& item_err, std::string& err) = 0;
    // Waits for the step issued into set w; its completion status.
    virtual int step_wait(int w, std::string& err) = 0;
};

struct Options {
    int max_slots = 16;
    int64_t budget = 64ll << 30;
    int edge_mode = PMP_EDGE_WRAP;
    int inflight_kpx = 0;
    int tail_stages = 0;
};

class Server {
  public:
    // Starts the scheduler thread.  The executor outlives the server.
    Server(Executor* ex, const Options& o);
    // Finishes every queued request (unless paused), then stops the thread.
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    int submit(const pmp_request& req, const pmp_params& p, uint64_t* ticket, std::string& err);
    int wait(uint64_t ticket, int timeout_ms, std::string& err);
    int fetch(uint64_t ticket, float* plane4, float* depth, float* score, int32_t* nghbr, pmp_request_info* info,
              std::string& err);
    int cancel(uint64_t ticket, std::string& err);
    int reset(std::string& err);
    int pause(bool paused);
    int stats(pmp_server_info* out);

  private:
    enum { kQueued = 0, kActive = 1, kDone = 2, kFailed = 3, kCancelled = 4 };
    struct HostBuf {
        uint8_t* p = nullptr;
        size_t cap = 0;
    };
    struct Req {
        Job job;
        int state = kQueued;
        int slot = -1;
        int stage = 0;
        int steps = 0;
        int inflight = 0;      // issued steps that include it and are not collected yet
        bool ending = false;   // no further stage: ends when inflight reaches 0
        int rc = PMP_OK;
        std::string error;
        double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
        std::vector<uint64_t> pins;
        HostBuf out, lm, prev;
    };
    struct Entry {
        int W = 0, H = 0;
        bool packed = false;
        int pins = 0;
        uint64_t last_use = 0;
        int64_t bytes = 0;
    };

    Executor* ex_;
    Options o_;
    std::chrono::steady_clock::time_point t0_;
    std::mutex mu_;          // all state below
    std::mutex submit_mu_;   // one submission (upload) or reset at a time
    std::condition_variable cv_work_, cv_done_, cv_room_;
    std::thread th_;
    bool stop_ = false, paused_ = false;

    std::deque<Req*> queue_;
    std::unordered_map<uint64_t, Req*> reqs_;
    uint64_t next_ticket_ = 1;
    std::vector<Req*> slot_req_;
    std::vector<int64_t> slot_cap_;   // pixels of the slot's maps
    std::vector<HostBuf> free_host_;
    std::unordered_map<uint64_t, Entry> pool_;
    uint64_t clock_ = 0;
    int64_t used_ = 0;                // pooled image bytes + slot map bytes + reservations
    std::vector<StepItem> step_items_[2];
    bool step_pending_[2] = {false, false};

    int64_t submitted_ = 0, completed_ = 0, failed_ = 0, cancelled_ = 0, steps_ = 0, uploads_ = 0, evictions_ = 0;
    int max_in_flight_ = 0;

    double now_ms() const;
    void loop();
    bool issue_step(int w);
    void collect(int w);
    void admit();
    void end_request(Req* r, int rc, const std::string& msg);
    void retire(Req* r);
    bool make_room(int64_t need);
    int reserve(std::unique_lock<std::mutex>& lk, int64_t need, std::string& err);
    int acquire(uint64_t id, const uint8_t* pixels, int W, int H, bool pack, Req* r, std::string& err);
    void unpin(Req* r);
    HostBuf take_host(size_t bytes);
    void give_host(HostBuf& b);
    bool busy_locked() const;
};

}  // namespace sched
}  // namespace pmp
