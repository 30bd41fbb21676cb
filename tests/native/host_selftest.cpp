// host_selftest.cpp -- the host-only parts of libpmp under AddressSanitizer +
// UndefinedBehaviorSanitizer (and the scheduler under ThreadSanitizer), on
// the CPU, with no device:
//   * the serving scheduler (pmp_sched.cpp) against a simulated executor
//     that checks every device-side invariant the scheduler promises (no
//     step reads an evicted image or an undersized slot, memory stays within
//     the budget, results are written by a step before anyone may read
//     them, stages run 0..last in order), with tiny budgets, cancellation,
//     reset, pause, and injected failures of uploads, packing, slot
//     allocation and steps (at issue and at completion);
//   * the dense-map writers (pmp_io.hip) and the densify-thread front-end's
//     parallax gate and depth interval (pmp_frontend.hip).
// Exits non-zero on the first failed check (or a sanitizer report).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <map>
#include <mutex>
#include <random>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "pmp/pmp.h"
#include "pmp_sched.h"

using namespace pmp;

#define CHECK(c)                                                                  \
    do {                                                                          \
        if (!(c)) {                                                               \
            fprintf(stderr, "host_selftest: check failed at line %d: %s\n", __LINE__, #c); \
            exit(1);                                                              \
        }                                                                         \
    } while (0)

// ---------------------------------------------------------------- simulated executor
struct SimExec final : sched::Executor {
    std::mutex mu;
    int64_t budget;
    int64_t bytes = 0;                       // images + slot maps held
    struct Img {
        int W, H;
        bool packed;
    };
    std::map<uint64_t, Img> images;
    std::vector<int64_t> slot_pixels;
    // deferred "device" work of descriptor sets 0 and 1: result writes
    struct Write {
        uint8_t* out;
        size_t n;
        uint32_t seed;
    };
    std::vector<Write> pending[2];
    std::set<uint64_t> reading[2];           // images read by the step in flight in each set
    std::map<uint64_t, std::vector<int>> stages;   // ticket -> stages run, in order
    std::map<std::pair<uint64_t, int>, int> step_of;   // (ticket, stage) -> index of the step that issued it
    std::map<uint64_t, std::vector<double>> prev_seen;  // ticket -> warm-start values read at stage 0
    std::atomic<int> issues{0}, waits{0};
    // fault injection
    uint64_t fail_upload_id = ~0ull, fail_pack_id = ~0ull;
    int fail_slot_grow_at = -1, slot_grows = 0;
    int fail_issue_at = -1, fail_wait_at = -1;
    int fail_item_view = -1;                 // step items of requests with this many views fail
    int step_us = 0;                         // simulated step duration
    std::atomic<int> violations{0};

    SimExec(int64_t b, int slots) : budget(b), slot_pixels(slots, 0) {}

    void bad(const char* what) {
        fprintf(stderr, "host_selftest: executor invariant broken: %s\n", what);
        violations++;
    }
    int check_params(const pmp_params& p, std::string& err) override {
        if (p.iterations < 0 || p.iterations > 64) {
            err = "iterations";
            return PMP_EINVAL;
        }
        return PMP_OK;
    }
    int64_t image_bytes(int W, int H, bool packed) const override { return (int64_t)W * H * (packed ? 3 : 1); }
    int64_t map_bytes(int64_t pixels) const override { return pixels * 84; }
    int upload(uint64_t id, const uint8_t* px, int W, int H, std::string& err) override {
        std::lock_guard<std::mutex> lk(mu);
        if (id == fail_upload_id) {
            err = "injected upload failure";
            return PMP_EDEVICE;
        }
        volatile uint8_t sink = px[(size_t)W * H - 1];   // the host pixels must be readable
        (void)sink;
        if (images.count(id)) bad("upload of a pooled id");
        images[id] = Img{W, H, false};
        bytes += image_bytes(W, H, false);
        if (bytes > budget) bad("upload beyond the budget");
        return PMP_OK;
    }
    int pack(uint64_t id, int wrap, std::string& err) override {
        (void)wrap;
        std::lock_guard<std::mutex> lk(mu);
        if (id == fail_pack_id) {
            err = "injected packing failure";
            return PMP_EDEVICE;
        }
        auto it = images.find(id);
        if (it == images.end()) {
            bad("packing an absent image");
            return PMP_ENOTFOUND;
        }
        if (it->second.packed) bad("packing twice");
        it->second.packed = true;
        bytes += image_bytes(it->second.W, it->second.H, true) - image_bytes(it->second.W, it->second.H, false);
        if (bytes > budget) bad("packing beyond the budget");
        return PMP_OK;
    }
    void drop(uint64_t id) override {
        std::lock_guard<std::mutex> lk(mu);
        auto it = images.find(id);
        if (it == images.end()) return;
        if (reading[0].count(id) || reading[1].count(id)) bad("dropping an image a running step reads");
        bytes -= image_bytes(it->second.W, it->second.H, it->second.packed);
        images.erase(it);
    }
    int slot_grow(int k, int64_t pixels, std::string& err) override {
        std::lock_guard<std::mutex> lk(mu);
        bytes -= map_bytes(slot_pixels[k]);
        slot_pixels[k] = 0;
        if (slot_grows++ == fail_slot_grow_at) {
            err = "injected allocation failure";
            return PMP_ENOMEM;
        }
        slot_pixels[k] = pixels;
        bytes += map_bytes(pixels);
        if (bytes > budget) bad("slot maps beyond the budget");
        return PMP_OK;
    }
    uint8_t* host_alloc(size_t n) override { return (uint8_t*)malloc(n); }
    void host_free(uint8_t* p) override { free(p); }

    int step_issue(int w, const std::vector<sched::StepItem>& items, std::vector<int>& item_rc,
                   std::vector<std::string>& item_err, std::string& err) override {
        std::lock_guard<std::mutex> lk(mu);
        const int n = issues++;
        if (!pending[w].empty() || !reading[w].empty()) bad("descriptor set reused before its step was collected");
        if (n == fail_issue_at) {
            err = "injected launch failure";
            return PMP_EDEVICE;
        }
        for (size_t k = 0; k < items.size(); ++k) {
            const sched::StepItem& it = items[k];
            const sched::Job* j = it.job;
            if (j->kf.n_views == fail_item_view) {
                item_rc[k] = PMP_EINVAL;
                item_err[k] = "injected descriptor failure";
                continue;
            }
            if (slot_pixels[it.slot] < (int64_t)j->W * j->H) bad("slot smaller than the keyframe");
            auto need = [&](uint64_t id, bool packed) {
                auto im = images.find(id);
                if (im == images.end()) bad("step reads an absent image");
                else if (im->second.W != j->W || im->second.H != j->H) bad("image size");
                else if (packed && !im->second.packed) bad("neighbour not packed");
                reading[w].insert(id);
            };
            need(j->kf.ref_id, false);
            for (int v = 0; v < j->kf.n_views; ++v) need(j->kf.view_ids[v], true);
            if (j->kf.mask_id >= 0) need((uint64_t)j->kf.mask_id, false);
            auto& st = stages[j->ticket];
            if ((int)st.size() != it.stage) bad("stage out of order");
            st.push_back(it.stage);
            step_of[{j->ticket, it.stage}] = n;
            if (it.stage == 0 && j->n_lm > 0) {
                volatile double sink = j->lm[3 * j->n_lm - 1];   // the landmarks must be readable
                (void)sink;
            }
            if (it.stage == 0 && j->prev_plane4) {
                // the previous result must be readable in full at the warm start
                const size_t P = (size_t)j->W * j->H;
                if (!j->prev_depth) bad("warm start without previous depth");
                else prev_seen[j->ticket] = {j->prev_plane4[0], j->prev_plane4[4 * P - 1], j->prev_depth[0],
                                             j->prev_depth[P - 1], j->prev_pose[0], j->prev_pose[11]};
            }
            if (it.stage == j->last_stage) {
                const size_t P = (size_t)j->W * j->H;
                pending[w].push_back(Write{j->out, P * 28, (uint32_t)(j->kf.ref_id * 2654435761u) ^ (uint32_t)j->p.seed});
            }
        }
        return PMP_OK;
    }
    int step_wait(int w, std::string& err) override {
        if (step_us) std::this_thread::sleep_for(std::chrono::microseconds(step_us));
        std::lock_guard<std::mutex> lk(mu);
        const int n = waits++;
        reading[w].clear();
        if (n == fail_wait_at) {
            pending[w].clear();
            err = "injected step failure";
            return PMP_EDEVICE;
        }
        for (const Write& wr : pending[w])
            for (size_t b = 0; b < wr.n; ++b) wr.out[b] = (uint8_t)(wr.seed + b * 7);
        pending[w].clear();
        return PMP_OK;
    }
};

// expected depth word k of a request (the simulated device's result)
static bool result_ok(const std::vector<float>& depth, uint64_t ref_id, uint64_t seed, size_t P) {
    const uint32_t s = (uint32_t)(ref_id * 2654435761u) ^ (uint32_t)seed;
    const uint8_t* b = (const uint8_t*)depth.data();
    for (size_t k = 0; k < P * 4; ++k)
        if (b[k] != (uint8_t)(s + (P * 16 + k) * 7)) return false;
    return true;
}

struct Images {
    std::map<uint64_t, std::vector<uint8_t>> px;
    const uint8_t* get(uint64_t id, int W, int H) {
        auto& v = px[id];
        v.assign((size_t)W * H, (uint8_t)id);
        return v.data();
    }
};

static pmp_params params(int iterations, uint64_t seed) {
    pmp_params p;
    memset(&p, 0, sizeof(p));
    p.iterations = iterations;
    p.seed = seed;
    p.edge_mode = PMP_EDGE_WRAP;
    return p;
}

// A request over pooled ids: reference id, n views from `view0`, optional mask.
static pmp_request request(Images& im, uint64_t ref, uint64_t view0, int n, int W, int H, int64_t mask = -1,
                           bool with_pixels = true) {
    pmp_request r;
    memset(&r, 0, sizeof(r));
    r.kf.ref_id = ref;
    r.kf.n_views = n;
    for (int v = 0; v < n; ++v) r.kf.view_ids[v] = view0 + v;
    r.kf.mask_id = mask;
    r.width = W;
    r.height = H;
    if (with_pixels) {
        r.ref_pixels = im.get(ref, W, H);
        for (int v = 0; v < n; ++v) r.view_pixels[v] = im.get(view0 + v, W, H);
        if (mask >= 0) r.mask_pixels = im.get((uint64_t)mask, W, H);
    }
    return r;
}

static int fetch(sched::Server& s, uint64_t t, int W, int H, std::vector<float>* depth = nullptr) {
    std::string err;
    std::vector<float> d((size_t)W * H);
    CHECK(s.wait(t, -1, err) == PMP_OK);
    pmp_request_info info;
    const int rc = s.fetch(t, nullptr, d.data(), nullptr, nullptr, &info, err);
    if (depth) *depth = d;
    return rc;
}

static void stages_complete(SimExec& ex, uint64_t t, int iterations) {
    std::lock_guard<std::mutex> lk(ex.mu);
    const auto& st = ex.stages[t];
    CHECK((int)st.size() == 2 * iterations + 2);
    for (size_t k = 0; k < st.size(); ++k) CHECK(st[k] == (int)k);
}

// 1. A stream from two threads under a budget that forces evictions.
static void test_stream() {
    const int W = 40, H = 20;
    const int64_t budget = 84ll * W * H * 3 + 3ll * W * H * 12;   // 3 slots of maps, about 12 packed images
    SimExec ex(budget, 3);
    ex.step_us = 50;
    sched::Options o;
    o.max_slots = 3;
    o.budget = budget;
    std::vector<std::pair<uint64_t, uint64_t>> tickets[2];   // (ticket, ref id)
    {
        sched::Server s(&ex, o);
        std::vector<std::thread> th;
        for (int t = 0; t < 2; ++t)
            th.emplace_back([&, t] {
                Images im;
                std::mt19937 g(7 + t);
                for (int k = 0; k < 30; ++k) {
                    const uint64_t ref = 100 * t + g() % 20;
                    const uint64_t view0 = 1000 * (t + 1) + g() % 30;
                    const int n = 1 + g() % 3;
                    pmp_request r = request(im, ref, view0, n, W, H, (k % 4 == 0) ? (int64_t)(5000 + t) : -1);
                    pmp_params p = params(1 + k % 3, ref);
                    uint64_t tk;
                    std::string err;
                    const int rc = s.submit(r, p, &tk, err);
                    CHECK(rc == PMP_OK);
                    tickets[t].push_back({tk, ref});
                }
            });
        for (auto& t : th) t.join();
        for (int t = 0; t < 2; ++t)
            for (auto& tk : tickets[t]) {
                std::vector<float> d;
                CHECK(fetch(s, tk.first, W, H, &d) == PMP_OK);
                CHECK(result_ok(d, tk.second, tk.second, (size_t)W * H));
            }
        pmp_server_info info;
        s.stats(&info);
        CHECK(info.completed == 60 && info.failed == 0 && info.in_flight == 0 && info.queued == 0);
        CHECK(info.evictions > 0);
        CHECK(info.max_in_flight >= 2);
    }
    CHECK(ex.violations == 0);
    for (int t = 0; t < 2; ++t)
        for (auto& tk : tickets[t]) {
            std::lock_guard<std::mutex> lk(ex.mu);
            const auto& st = ex.stages[tk.first];
            CHECK(st.size() >= 4 && st.size() % 2 == 0);
            for (size_t k = 0; k < st.size(); ++k) CHECK(st[k] == (int)k);
        }
}

// 2. Cancel while queued and while in flight; the slot takes the next request.
static void test_cancel() {
    const int W = 32, H = 16;
    SimExec ex(1ll << 30, 1);
    ex.step_us = 200;
    sched::Options o;
    o.max_slots = 1;
    o.budget = 1ll << 30;
    Images im;
    sched::Server s(&ex, o);
    std::string err;
    s.pause(true);
    uint64_t t[5];
    for (int k = 0; k < 5; ++k) {
        pmp_request r = request(im, 10 + k, 20, 2, W, H);
        CHECK(s.submit(r, params(k == 0 ? 40 : 2, 10 + k), &t[k], err) == PMP_OK);
    }
    CHECK(s.cancel(t[3], err) == PMP_OK);          // queued
    s.pause(false);
    // t[0] runs 82 stages of >= 200 us: cancel it in flight
    for (int spin = 0; spin < 2000; ++spin) {
        pmp_server_info info;
        s.stats(&info);
        if (info.in_flight == 1 && info.steps > 3) break;
        std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    CHECK(s.cancel(t[0], err) == PMP_OK);
    CHECK(fetch(s, t[0], W, H) == PMP_ECANCELED);
    CHECK(fetch(s, t[3], W, H) == PMP_ECANCELED);
    for (int k : {1, 2, 4}) {
        std::vector<float> d;
        CHECK(fetch(s, t[k], W, H, &d) == PMP_OK);
        CHECK(result_ok(d, 10 + k, 10 + k, (size_t)W * H));
        stages_complete(ex, t[k], 2);
    }
    {
        std::lock_guard<std::mutex> lk(ex.mu);
        CHECK(ex.stages[t[0]].size() < 82);         // stopped early
        CHECK(ex.stages.count(t[3]) == 0);          // never ran
    }
    CHECK(s.cancel(12345, err) == PMP_ENOTFOUND);
    pmp_server_info info;
    s.stats(&info);
    CHECK(info.cancelled == 2 && info.completed == 3);
    CHECK(ex.violations == 0);
}

// 3. Reset: everything queued or in flight ends cancelled, the pool empties,
// the server keeps serving.
static void test_reset() {
    const int W = 32, H = 16;
    SimExec ex(1ll << 30, 2);
    ex.step_us = 200;
    sched::Options o;
    o.max_slots = 2;
    o.budget = 1ll << 30;
    Images im;
    sched::Server s(&ex, o);
    std::string err;
    std::vector<uint64_t> t(8);
    for (int k = 0; k < 8; ++k) {
        pmp_request r = request(im, 30 + k, 40 + k, 2, W, H);
        CHECK(s.submit(r, params(20, k), &t[k], err) == PMP_OK);
    }
    std::this_thread::sleep_for(std::chrono::milliseconds(2));
    CHECK(s.reset(err) == PMP_OK);
    pmp_server_info info;
    s.stats(&info);
    CHECK(info.in_flight == 0 && info.queued == 0 && info.pool_images == 0 && info.pool_bytes == 0);
    int cancelled = 0;
    for (int k = 0; k < 8; ++k) {
        const int rc = fetch(s, t[k], W, H);
        CHECK(rc == PMP_ECANCELED || rc == PMP_OK);
        cancelled += rc == PMP_ECANCELED;
    }
    CHECK(cancelled >= 6);
    {
        std::lock_guard<std::mutex> lk(ex.mu);
        CHECK(ex.images.empty());
    }
    // pooled ids are gone: a request without pixels fails, one with pixels runs
    pmp_request bare = request(im, 30, 40, 2, W, H, -1, false);
    uint64_t tk;
    CHECK(s.submit(bare, params(1, 0), &tk, err) == PMP_ENOTFOUND);
    pmp_request full = request(im, 30, 40, 2, W, H);
    CHECK(s.submit(full, params(1, 30), &tk, err) == PMP_OK);
    std::vector<float> d;
    CHECK(fetch(s, tk, W, H, &d) == PMP_OK);
    CHECK(result_ok(d, 30, 30, (size_t)W * H));
    CHECK(ex.violations == 0);
}

// 4. Injected failures: each fails only the requests it concerns, releases
// what they held, and the server keeps serving.
static void test_faults() {
    const int W = 24, H = 12;
    const size_t P = (size_t)W * H;
    std::string err;
    {   // upload and packing failures at submission
        SimExec ex(1ll << 30, 2);
        ex.fail_upload_id = 77;
        ex.fail_pack_id = 88;
        sched::Options o;
        o.max_slots = 2;
        o.budget = 1ll << 30;
        Images im;
        sched::Server s(&ex, o);
        uint64_t tk;
        pmp_request r1 = request(im, 1, 76, 2, W, H);    // views 76, 77: 77 fails
        CHECK(s.submit(r1, params(1, 1), &tk, err) == PMP_EDEVICE);
        pmp_request r2 = request(im, 2, 87, 2, W, H);    // views 87, 88: 88's packing fails
        CHECK(s.submit(r2, params(1, 2), &tk, err) == PMP_EDEVICE);
        pmp_request r3 = request(im, 3, 76, 1, W, H);    // view 76 is pooled (pinned by nobody now)
        CHECK(s.submit(r3, params(1, 3), &tk, err) == PMP_OK);
        std::vector<float> d;
        CHECK(fetch(s, tk, W, H, &d) == PMP_OK && result_ok(d, 3, 3, P));
        pmp_server_info info;
        s.stats(&info);
        CHECK(info.failed == 0 && info.completed == 1);
        CHECK(ex.violations == 0);
    }
    {   // a step that cannot be issued fails everything in flight
        SimExec ex(1ll << 30, 3);
        ex.fail_issue_at = 4;
        sched::Options o;
        o.max_slots = 3;
        o.budget = 1ll << 30;
        Images im;
        sched::Server s(&ex, o);
        std::vector<uint64_t> t(6);
        s.pause(true);
        for (int k = 0; k < 6; ++k) {
            pmp_request r = request(im, 10 + k, 50, 2, W, H);
            CHECK(s.submit(r, params(3, 10 + k), &t[k], err) == PMP_OK);
        }
        s.pause(false);
        int failed = 0, ok = 0;
        for (int k = 0; k < 6; ++k) {
            std::vector<float> d;
            const int rc = fetch(s, t[k], W, H, &d);
            if (rc == PMP_OK) {
                CHECK(result_ok(d, 10 + k, 10 + k, P));
                ++ok;
            } else {
                CHECK(rc == PMP_EDEVICE);
                ++failed;
            }
        }
        CHECK(failed >= 1 && failed <= 3 && ok >= 3);
        CHECK(ex.violations == 0);
    }
    {   // a step that fails on the device fails the requests it carried
        SimExec ex(1ll << 30, 2);
        ex.fail_wait_at = 2;
        sched::Options o;
        o.max_slots = 2;
        o.budget = 1ll << 30;
        Images im;
        sched::Server s(&ex, o);
        std::vector<uint64_t> t(4);
        s.pause(true);
        for (int k = 0; k < 4; ++k) {
            pmp_request r = request(im, 10 + k, 60, 1, W, H);
            CHECK(s.submit(r, params(2, 10 + k), &t[k], err) == PMP_OK);
        }
        s.pause(false);
        int failed = 0;
        for (int k = 0; k < 4; ++k) {
            const int rc = fetch(s, t[k], W, H);
            CHECK(rc == PMP_OK || rc == PMP_EDEVICE);
            failed += rc == PMP_EDEVICE;
        }
        CHECK(failed >= 1 && failed <= 2);
        CHECK(ex.violations == 0);
    }
    {   // a descriptor failure fails its own request only
        SimExec ex(1ll << 30, 2);
        ex.fail_item_view = 3;
        sched::Options o;
        o.max_slots = 2;
        o.budget = 1ll << 30;
        Images im;
        sched::Server s(&ex, o);
        uint64_t a, b;
        pmp_request ra = request(im, 1, 70, 3, W, H);
        pmp_request rb = request(im, 2, 70, 2, W, H);
        CHECK(s.submit(ra, params(2, 1), &a, err) == PMP_OK);
        CHECK(s.submit(rb, params(2, 2), &b, err) == PMP_OK);
        CHECK(fetch(s, a, W, H) == PMP_EINVAL);
        std::vector<float> d;
        CHECK(fetch(s, b, W, H, &d) == PMP_OK && result_ok(d, 2, 2, P));
        CHECK(ex.violations == 0);
    }
    {   // slot maps that cannot be allocated; budgets too small
        SimExec ex(84ll * 64 * 32 + 3ll * 64 * 32 * 4, 1);
        ex.fail_slot_grow_at = 0;
        sched::Options o;
        o.max_slots = 1;
        o.budget = 84ll * 64 * 32 + 3ll * 64 * 32 * 4;
        Images im;
        sched::Server s(&ex, o);
        uint64_t tk;
        pmp_request r = request(im, 1, 2, 1, 64, 32);
        CHECK(s.submit(r, params(1, 1), &tk, err) == PMP_OK);
        CHECK(fetch(s, tk, 64, 32) == PMP_ENOMEM);   // injected allocation failure
        CHECK(s.submit(r, params(1, 1), &tk, err) == PMP_OK);
        std::vector<float> d;
        CHECK(fetch(s, tk, 64, 32, &d) == PMP_OK && result_ok(d, 1, 1, 64 * 32));
        pmp_request big = request(im, 9, 10, 1, 256, 128);
        CHECK(s.submit(big, params(1, 9), &tk, err) == PMP_ENOMEM);   // maps larger than the budget
        pmp_request many = request(im, 20, 21, 8, 64, 32);
        CHECK(s.submit(many, params(1, 20), &tk, err) == PMP_ENOMEM);   // images can never fit
        pmp_params bad = params(99, 0);
        CHECK(s.submit(r, bad, &tk, err) == PMP_EINVAL);
        pmp_params zero = params(1, 0);
        zero.edge_mode = PMP_EDGE_ZERO;
        CHECK(s.submit(r, zero, &tk, err) == PMP_EINVAL);
        CHECK(ex.violations == 0);
    }
}

// 5. Destroying a server finishes what was queued; landmarks travel with
// their request.
static void test_destroy_finishes() {
    const int W = 16, H = 8;
    SimExec ex(1ll << 30, 2);
    ex.step_us = 100;
    std::vector<uint64_t> t(5);
    {
        sched::Options o;
        o.max_slots = 2;
        o.budget = 1ll << 30;
        Images im;
        sched::Server s(&ex, o);
        std::string err;
        std::vector<double> lm(3 * 50, 1.5);
        for (int k = 0; k < 5; ++k) {
            pmp_request r = request(im, 1 + k, 8, 2, W, H);
            r.n_landmarks = 50;
            r.landmark_xyz = lm.data();
            CHECK(s.submit(r, params(2, k), &t[k], err) == PMP_OK);
        }
    }
    for (uint64_t tk : t) stages_complete(ex, tk, 2);
    CHECK(ex.violations == 0);
}

// 5b. A warm start travels with its request as a copy made at submit (the
// caller's maps may go right after), reaches the executor at stage 0, and is
// all or nothing.
static void test_warm_start() {
    const int W = 16, H = 8;
    const size_t P = (size_t)W * H;
    SimExec ex(1ll << 30, 2);
    uint64_t t0 = 0, t1 = 0;
    {
        sched::Options o;
        o.max_slots = 2;
        o.budget = 1ll << 30;
        Images im;
        sched::Server s(&ex, o);
        std::string err;
        std::vector<float> plane(4 * P), depth(P);
        for (size_t k = 0; k < 4 * P; ++k) plane[k] = 0.25f * (float)k;
        for (size_t k = 0; k < P; ++k) depth[k] = 1.0f + (float)k;
        pmp_pose pose;
        for (int k = 0; k < 9; ++k) pose.R[k] = k;
        for (int k = 0; k < 3; ++k) pose.t[k] = 9 + k;
        pmp_request r = request(im, 1, 8, 2, W, H);
        r.prev_plane4 = plane.data();
        r.prev_depth = depth.data();
        CHECK(s.submit(r, params(2, 0), &t0, err) == PMP_EINVAL);   // no pose: all or nothing
        r.prev_pose = &pose;
        CHECK(s.submit(r, params(2, 0), &t0, err) == PMP_OK);
        std::fill(plane.begin(), plane.end(), -1.0f);                  // the request holds its own copy
        std::fill(depth.begin(), depth.end(), -1.0f);
        pmp_request cold = request(im, 2, 8, 2, W, H);
        CHECK(s.submit(cold, params(2, 1), &t1, err) == PMP_OK);
        CHECK(s.wait(t0, -1, err) == PMP_OK && s.wait(t1, -1, err) == PMP_OK);
    }
    stages_complete(ex, t0, 2);
    stages_complete(ex, t1, 2);
    CHECK(ex.prev_seen.count(t0) == 1 && ex.prev_seen.count(t1) == 0);
    const std::vector<double>& v = ex.prev_seen[t0];
    CHECK(v[0] == 0.0 && v[1] == 0.25 * (4 * P - 1) && v[2] == 1.0 && v[3] == (double)P && v[4] == 0.0 && v[5] == 11.0);
    CHECK(ex.violations == 0);
}

// 6. Slot choice: a request takes the smallest free slot whose maps already
// hold it, else the largest free slot, so slot maps grow only when no free
// slot is large enough.
static void test_slot_best_fit() {
    const int Ws = 16, Hs = 8, Wb = 64, Hb = 32;
    SimExec ex(1ll << 30, 2);
    sched::Options o;
    o.max_slots = 2;
    o.budget = 1ll << 30;
    Images im;
    sched::Server s(&ex, o);
    std::string err;
    uint64_t a, b, c, d;
    s.pause(true);   // admit the small and the large request together
    CHECK(s.submit(request(im, 1, 10, 2, Ws, Hs), params(1, 1), &a, err) == PMP_OK);
    CHECK(s.submit(request(im, 2, 20, 2, Wb, Hb), params(1, 2), &b, err) == PMP_OK);
    s.pause(false);
    CHECK(fetch(s, a, Ws, Hs) == PMP_OK);
    CHECK(fetch(s, b, Wb, Hb) == PMP_OK);
    {
        std::lock_guard<std::mutex> lk(ex.mu);
        CHECK(ex.slot_grows == 2);
        CHECK(ex.slot_pixels[0] == (int64_t)Ws * Hs && ex.slot_pixels[1] == (int64_t)Wb * Hb);
    }
    // both slots free: the large request takes slot 1, the small one slot 0
    CHECK(s.submit(request(im, 3, 30, 2, Wb, Hb), params(1, 3), &c, err) == PMP_OK);
    CHECK(fetch(s, c, Wb, Hb) == PMP_OK);
    CHECK(s.submit(request(im, 4, 40, 2, Ws, Hs), params(1, 4), &d, err) == PMP_OK);
    CHECK(fetch(s, d, Ws, Hs) == PMP_OK);
    {
        std::lock_guard<std::mutex> lk(ex.mu);
        CHECK(ex.slot_grows == 2);
    }
    for (uint64_t t : {a, b, c, d}) stages_complete(ex, t, 1);
    CHECK(ex.violations == 0);
}

// 7. Tail admission: with one request queued, it joins a request in flight
// only for that request's last tail_stages stages; with a backlog of two or
// more, admission is free.  (3 iterations: stages 0..7.)
static void test_tail_admission() {
    const int W = 16, H = 8;
    for (int mode = 0; mode < 3; ++mode) {   // 0: tail 2; 1: tail off; 2: tail 2 with a backlog
        SimExec ex(1ll << 30, 2);
        ex.step_us = 1000;   // the second submit comes microseconds after the first
        sched::Options o;
        o.max_slots = 2;
        o.budget = 1ll << 30;
        o.tail_stages = mode == 1 ? 0 : 2;
        Images im;
        uint64_t a, b, c = 0;
        {
            sched::Server s(&ex, o);
            std::string err;
            if (mode == 2) s.pause(true);
            CHECK(s.submit(request(im, 1, 10, 2, W, H), params(3, 1), &a, err) == PMP_OK);
            CHECK(s.submit(request(im, 2, 20, 2, W, H), params(3, 2), &b, err) == PMP_OK);
            if (mode == 2) {
                CHECK(s.submit(request(im, 3, 30, 2, W, H), params(3, 3), &c, err) == PMP_OK);
                s.pause(false);
            }
            CHECK(fetch(s, a, W, H) == PMP_OK);
            CHECK(fetch(s, b, W, H) == PMP_OK);
            if (mode == 2) CHECK(fetch(s, c, W, H) == PMP_OK);
        }
        std::lock_guard<std::mutex> lk(ex.mu);
        const int b0 = ex.step_of.at({b, 0});
        if (mode == 0) CHECK(b0 == ex.step_of.at({a, 6}));   // joins for a's last 2 stages
        if (mode == 1) CHECK(b0 <= ex.step_of.at({a, 2}));   // joins at once
        if (mode == 2) CHECK(b0 == ex.step_of.at({a, 0}));   // backlog: joins at once
        CHECK(ex.violations == 0);
    }
}

// ---------------------------------------------------------------- host-only library code
static void test_frontend_and_writers() {
    // depth interval: 1e-5 and 1.1 x the distance at sorted index n/10*9
    double pose[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
    std::vector<double> xyz;
    for (int k = 1; k <= 20; ++k) {
        xyz.push_back(0.0);
        xyz.push_back(0.0);
        xyz.push_back((double)k);
    }
    double lo = 0, hi = 0;
    CHECK(pmp_depth_interval(pose, 20, xyz.data(), &lo, &hi) == PMP_OK);
    CHECK(lo == 1e-5 && fabs(hi - 1.1 * 19.0) < 1e-9);
    CHECK(pmp_depth_interval(pose, 0, xyz.data(), &lo, &hi) == PMP_EINVAL);
    // parallax gate: landmarks 3 m away, cameras 0.5 m apart -> angles inside (pi/60, pi/6)
    std::vector<uint64_t> ids;
    std::vector<double> pts;
    std::mt19937 g(3);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (int k = 0; k < 200; ++k) {
        ids.push_back(1000 + k);
        pts.push_back(u(g));
        pts.push_back(u(g));
        pts.push_back(3.0 + u(g));
    }
    double p2[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, -0.5, 0, 0};
    const double sv = pmp_score_view(pose, 200, ids.data(), pts.data(), p2, 200, ids.data(), pts.data());
    CHECK(sv > 0.9 && sv <= 1.0);
    std::vector<uint64_t> other(ids.size());
    for (size_t k = 0; k < ids.size(); ++k) other[k] = ids[k] + 5000;
    CHECK(isnan(pmp_score_view(pose, 200, ids.data(), pts.data(), p2, 200, other.data(), pts.data())));
    // writers
    const char* ply = "host_selftest.ply";
    std::vector<uint8_t> rgb(3 * 200, 77);
    CHECK(pmp_write_ply(ply, 200, pts.data(), rgb.data()) == PMP_OK);
    FILE* f = fopen(ply, "r");
    CHECK(f);
    char line[256];
    int lines = 0;
    while (fgets(line, sizeof(line), f)) ++lines;
    fclose(f);
    remove(ply);
    CHECK(lines == 200 + 10);
    pmp_camera_info cam{"equirect", 1920, 960, 30.0};
    std::vector<int32_t> kps(2 * 200, 3);
    pmp_dense_keyframe kf;
    memset(&kf, 0, sizeof(kf));
    kf.id = 7;
    kf.src_frm_id = 7;
    kf.pose.R[0] = kf.pose.R[4] = kf.pose.R[8] = 1.0;
    kf.n_points = 200;
    kf.points = pts.data();
    kf.colors = rgb.data();
    kf.keypts = kps.data();
    const char* mp = "host_selftest.msg";
    CHECK(pmp_write_dense_msgpack(mp, &cam, 1, &kf) == PMP_OK);
    f = fopen(mp, "rb");
    CHECK(f);
    fseek(f, 0, SEEK_END);
    const long sz = ftell(f);
    fclose(f);
    remove(mp);
    CHECK(sz > 200 * 20);
    CHECK(pmp_write_ply(nullptr, 1, pts.data(), rgb.data()) == PMP_EINVAL);
    CHECK(pmp_write_dense_msgpack(mp, &cam, -1, &kf) == PMP_EINVAL);
}

int main(int argc, char** argv) {
    const bool sched_only = argc > 1 && strcmp(argv[1], "--sched-only") == 0;
    test_stream();
    test_cancel();
    test_reset();
    test_faults();
    test_destroy_finishes();
    test_warm_start();
    test_slot_best_fit();
    test_tail_admission();
    if (!sched_only) test_frontend_and_writers();
    printf("host_selftest: ok\n");
    return 0;
}
