// pmp_server.hip -- the served form: a request stream densified with
// continuous batching (host code; the kernels are the work-queue kernels of
// pmp_queue.hip).
//
// A scheduler thread loops over steps.  A step admits queued requests into
// free slots (keyframe maps sized from the HBM budget), then launches one
// stage of every slot in flight, each at its own stage:
//   stage 0                      random init + initial score   (init list),
//                                then landmark seeding of the requests that
//                                carry landmarks (one launch pair each)
//   stage 1 .. 2*iterations      red-black pass stage-1        (propagation list)
//   stage 2*iterations + 1       copy-back, gate, PostProcess, copy to host
// so a request that arrives while others are on the device joins at the next
// step and its passes share launches with theirs.  Steps are issued two deep
// on the compute stream (descriptor sets alternate) so the device does not
// wait for the host between steps; the completions of step k are collected
// after step k+1 is queued.  Slot maps are independent, so every request's
// maps are those of pmp_densify with the same params.
//
// Images live in the context's pool, keyed by the caller's ids, with an LRU
// clock and a pin count (requests queued or in flight that read them);
// unpinned images are evicted oldest-first when uploads or slot maps need
// room in the budget.  Uploads and packing run on a second stream from the
// submitting thread and overlap the compute stream.
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "pmp_internal.h"
#include "pmp_queue.h"

namespace pmp {

namespace {

enum { kQueued = 0, kActive = 1, kDone = 2, kFailed = 3 };

struct HostBuf {
    uint8_t* p = nullptr;     // pinned
    size_t cap = 0;
};

struct Request {
    uint64_t ticket = 0;
    pmp_keyframe kf;
    pmp_params p;
    int W = 0, H = 0;
    int state = kQueued;
    int slot = -1;
    int stage = 0;
    int last_stage = 0;       // 2 * iterations + 1
    int steps = 0;
    int rc = PMP_OK;
    std::string error;
    double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
    std::vector<uint64_t> pins;
    HostBuf out;              // plane (16 B) | depth | score | nghbr per pixel
    HostBuf lm;               // landmark positions, 3 doubles each (pinned)
    int64_t n_lm = 0;
};

}  // namespace

}  // namespace pmp

using namespace pmp;

struct pmp_server {
    pmp_ctx* ctx = nullptr;
    hipStream_t copy = nullptr;
    pmp_server_opts opts;
    int max_slots = 16;
    int64_t budget = 0;
    int64_t used = 0;                    // pooled image bytes + slot map bytes
    std::chrono::steady_clock::time_point t0;

    std::mutex mu;
    std::mutex submit_mu;                // one submission (upload) at a time
    std::condition_variable cv_work, cv_done, cv_room;
    std::thread th;
    bool stop = false;
    bool paused = false;
    std::string last_error;

    std::deque<Request*> queue;
    std::unordered_map<uint64_t, Request*> reqs;
    uint64_t next_ticket = 1;
    std::vector<Request*> slot_req;
    std::vector<SlotBuf> slot_buf;
    std::vector<HostBuf> free_host;

    DescSet desc[2];
    hipEvent_t step_done[2] = {nullptr, nullptr};
    std::vector<Request*> finishing[2];   // requests whose finalize ran in that step
    bool step_pending[2] = {false, false};

    int64_t submitted = 0, completed = 0, failed = 0, steps = 0, uploads = 0, evictions = 0;
    int max_in_flight = 0;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace pmp {
namespace {

int srv_error(pmp_server* s, int code, const std::string& msg) {
    s->last_error = msg;
    return code;
}

size_t image_bytes(const Image& im) { return im.pool_bytes(); }

// Evicts unpinned pool images, least recently used first, until `need` more
// bytes fit in the budget.  Caller holds mu.  Returns false if they cannot.
bool make_room(pmp_server* s, int64_t need) {
    pmp_ctx* ctx = s->ctx;
    while (s->used + need > s->budget) {
        uint64_t victim = 0;
        uint64_t oldest = ~0ull;
        bool found = false;
        for (auto& kv : ctx->images)
            if (kv.second.pins == 0 && kv.second.last_use < oldest) {
                oldest = kv.second.last_use;
                victim = kv.first;
                found = true;
            }
        if (!found) return false;
        Image& im = ctx->images[victim];
        s->used -= (int64_t)image_bytes(im);
        // queued work never reads an unpinned image
        if (im.gray) (void)hipFree(im.gray);
        if (im.quads) (void)hipFree(im.quads);
        ctx->images.erase(victim);
        s->evictions++;
    }
    return true;
}

void unpin(pmp_server* s, Request* r) {
    for (uint64_t id : r->pins) {
        auto it = s->ctx->images.find(id);
        if (it != s->ctx->images.end() && it->second.pins > 0) it->second.pins--;
    }
    r->pins.clear();
}

HostBuf take_host(pmp_server* s, size_t bytes) {
    for (size_t k = 0; k < s->free_host.size(); ++k)
        if (s->free_host[k].cap >= bytes) {
            HostBuf b = s->free_host[k];
            s->free_host.erase(s->free_host.begin() + (long)k);
            return b;
        }
    HostBuf b;
    if (hipHostMalloc((void**)&b.p, bytes, hipHostMallocDefault) == hipSuccess) b.cap = bytes;
    else b.p = nullptr;
    return b;
}

void give_host(pmp_server* s, HostBuf& b) {
    if (b.p) s->free_host.push_back(b);
    b = HostBuf();
}

// Admission of queued requests into free slots, oldest first, within the
// in-flight pixel budget (caller holds mu).
void admit(pmp_server* s) {
    while (!s->queue.empty() && !s->paused) {
        int slot = -1;
        int64_t inflight = 0;
        for (int k = 0; k < s->max_slots; ++k) {
            if (s->slot_req[k]) inflight += (int64_t)s->slot_req[k]->W * s->slot_req[k]->H;
            else if (slot < 0) slot = k;
        }
        if (slot < 0) return;
        Request* r = s->queue.front();
        SlotBuf& b = s->slot_buf[slot];
        const size_t P = (size_t)r->W * r->H;
        if (s->opts.inflight_kpx > 0 && inflight > 0 && inflight + (int64_t)P > 1000ll * s->opts.inflight_kpx) return;
        if (P > b.cap) {
            const int64_t grow = (int64_t)((P - b.cap) * SlotBuf::kBytesPerPixel);
            if (!make_room(s, grow)) {
                bool flying = false;
                for (int k = 0; k < s->max_slots; ++k) flying = flying || s->slot_req[k];
                if (flying) return;   // wait for requests in flight to finish
                // nothing will free memory: the pinned images of queued requests leave no room
                s->queue.pop_front();
                r->state = kFailed;
                r->rc = PMP_ENOMEM;
                r->error = "HBM budget cannot hold the keyframe maps next to the queued requests' images";
                r->t_done = s->now_ms();
                unpin(s, r);
                s->failed++;
                s->cv_done.notify_all();
                s->cv_room.notify_all();
                continue;
            }
            s->used -= (int64_t)(b.cap * SlotBuf::kBytesPerPixel);
            free_slot(b);
            if (ensure_slot(s->ctx, b, r->W, r->H) != PMP_OK) {
                s->queue.pop_front();
                r->state = kFailed;
                r->rc = PMP_ENOMEM;
                r->error = "keyframe maps: " + s->ctx->last_error;
                r->t_done = s->now_ms();
                unpin(s, r);
                s->failed++;
                s->cv_done.notify_all();
                continue;
            }
            s->used += (int64_t)(P * SlotBuf::kBytesPerPixel);
        } else {
            b.W = r->W;
            b.H = r->H;
        }
        if (!r->out.p || r->out.cap < P * 28) {
            give_host(s, r->out);
            r->out = take_host(s, P * 28);
        }
        s->queue.pop_front();
        if (!r->out.p) {
            r->state = kFailed;
            r->rc = PMP_ENOMEM;
            r->error = "pinned host result buffer";
            r->t_done = s->now_ms();
            unpin(s, r);
            s->failed++;
            s->cv_done.notify_all();
            continue;
        }
        r->state = kActive;
        r->slot = slot;
        r->stage = 0;
        r->t_admit = s->now_ms();
        s->slot_req[slot] = r;
    }
}

void fail_active(pmp_server* s, Request* r, int rc, const std::string& msg) {
    r->state = kFailed;
    r->rc = rc;
    r->error = msg;
    r->t_done = s->now_ms();
    if (r->slot >= 0) s->slot_req[r->slot] = nullptr;
    r->slot = -1;
    unpin(s, r);
    s->failed++;
    s->cv_done.notify_all();
    s->cv_room.notify_all();
}

// Completions of the step issued into descriptor set w.
void collect(pmp_server* s, int w) {
    if (!s->step_pending[w]) return;
    const hipError_t e = hipEventSynchronize(s->step_done[w]);
    std::lock_guard<std::mutex> lk(s->mu);
    s->step_pending[w] = false;
    for (Request* r : s->finishing[w]) {
        if (e != hipSuccess) {
            fail_active(s, r, PMP_EDEVICE, hipGetErrorString(e));
            continue;
        }
        r->state = kDone;
        r->t_done = s->now_ms();
        s->slot_req[r->slot] = nullptr;
        r->slot = -1;
        unpin(s, r);
        s->completed++;
    }
    s->finishing[w].clear();
    s->cv_done.notify_all();
    s->cv_room.notify_all();
}

// Builds and issues one step into descriptor set w.  Returns false when
// nothing is in flight.
bool issue_step(pmp_server* s, int w) {
    pmp_ctx* ctx = s->ctx;
    DescSet& d = s->desc[w];
    int items_init[kMaxSlots] = {0}, items_prop[kMaxSlots] = {0}, items_rows[kMaxSlots] = {0};
    bool any = false, post = false;
    std::vector<Request*> fin;
    std::vector<std::pair<int, Request*>> seeds;   // slots whose init runs in this step and carry landmarks
    {
        std::lock_guard<std::mutex> lk(s->mu);
        admit(s);
        if (desc_begin(ctx, d) != PMP_OK) return false;
        int active = 0;
        for (int k = 0; k < s->max_slots; ++k) {
            Request* r = s->slot_req[k];
            if (!r) continue;
            // a request whose finalize is in the step before this one has no stage left
            if (r->stage > r->last_stage) continue;
            int items[3];
            const int pass0 = r->stage >= 1 && r->stage < r->last_stage ? r->stage - 1 : 0;
            int rc = slot_fill(ctx, &r->p, r->kf, s->slot_buf[k], nullptr, pass0, s->opts.config.tile_rows,
                               ctx->stream, d.h_slots[k], items);
            if (rc) {
                fail_active(s, r, rc, ctx->last_error);
                continue;
            }
            ++active;
            any = true;
            r->steps++;
            if (r->stage == 0) {
                items_init[k] = items[0];
                if (r->n_lm > 0) seeds.emplace_back(k, r);
            }
            else if (r->stage < r->last_stage) items_prop[k] = items[1];
            else {
                items_rows[k] = items[2];
                post = post || r->p.postprocess;
                fin.push_back(r);
            }
            r->stage++;
        }
        if (active > s->max_in_flight) s->max_in_flight = active;
        if (!any) return false;
        s->steps++;
    }
    fill_list(d.h_hdr->init, s->max_slots, items_init);
    fill_list(d.h_hdr->prop, s->max_slots, items_prop);
    fill_list(d.h_hdr->rows, s->max_slots, items_rows);
    QueueState* q = ctx->queue;
    QueueRun run;
    int rc = queue_run(ctx, *q, s->opts.config, s->opts.edge_mode == PMP_EDGE_ZERO, post, run);
    if (!rc) rc = desc_upload(ctx, d, s->max_slots, ctx->stream);
    if (!rc) rc = q_clear(ctx, ctx->stream, d, 1);
    if (!rc && d.h_hdr->init.n) rc = q_init(ctx, ctx->stream, d, run);
    for (const auto& kr : seeds) {
        if (rc) break;
        const Request* r = kr.second;   // alive until fetched, which needs this step's completion
        rc = preinit_landmarks_slot(ctx, d.d_slots + kr.first, r->W, r->H, (const double*)&r->kf.ref_pose, r->n_lm,
                                    (const double*)r->lm.p, s->opts.edge_mode == PMP_EDGE_ZERO);
    }
    if (!rc && d.h_hdr->prop.n) rc = q_prop(ctx, ctx->stream, d, run, 0);
    if (!rc && d.h_hdr->rows.n) rc = q_finalize(ctx, ctx->stream, d, run);
    for (Request* r : fin) {
        if (rc) break;
        const SlotBuf& b = s->slot_buf[r->slot];
        const size_t P = (size_t)r->W * r->H;
        uint8_t* o = r->out.p;
        hipError_t e = hipMemcpyAsync(o, b.oplane, P * 16, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(o + P * 16, b.odepth, P * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(o + P * 20, b.oscore, P * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(o + P * 24, b.onghbr, P * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e != hipSuccess) rc = check_hip(ctx, e, "result copy");
    }
    if (!rc && hipEventRecord(s->step_done[w], ctx->stream) != hipSuccess) rc = PMP_EDEVICE;
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc) {
        // a launch failed: fail everything in flight
        (void)hipStreamSynchronize(ctx->stream);
        const std::string msg = ctx->last_error;
        for (int k = 0; k < s->max_slots; ++k)
            if (s->slot_req[k]) fail_active(s, s->slot_req[k], rc, msg);
        return true;
    }
    s->finishing[w] = fin;
    s->step_pending[w] = true;
    return true;
}

void scheduler(pmp_server* s) {
    (void)hipSetDevice(s->ctx->device);
    int w = 0;
    for (;;) {
        const bool issued = issue_step(s, w);
        // collect the previous step while this one runs
        collect(s, w ^ 1);
        if (!issued) {
            collect(s, w);
            std::unique_lock<std::mutex> lk(s->mu);
            bool busy = false;
            for (int k = 0; k < s->max_slots; ++k) busy = busy || s->slot_req[k];
            if (busy) continue;
            if (s->stop && (s->queue.empty() || s->paused)) break;
            s->cv_work.wait(lk, [s] { return s->stop || (!s->queue.empty() && !s->paused); });
            continue;
        }
        w ^= 1;
    }
    collect(s, 0);
    collect(s, 1);
}

// Waits until `need` bytes fit in the budget (evicting unpinned images), then
// reserves them.  Caller holds lk on mu.
int reserve(pmp_server* s, std::unique_lock<std::mutex>& lk, int64_t need) {
    if (need > s->budget) return srv_error(s, PMP_ENOMEM, "image larger than the HBM budget");
    while (!make_room(s, need)) {
        bool busy = !s->queue.empty() && !s->paused;
        for (int k = 0; k < s->max_slots; ++k) busy = busy || s->slot_req[k];
        if (!busy) return srv_error(s, PMP_ENOMEM, "HBM budget cannot hold the request's images");
        s->cv_room.wait(lk);
    }
    s->used += need;
    return PMP_OK;
}

// Pins the pool entry of `id` for request r, uploading it from host pixels
// if absent and packing its column pairs if it serves as a neighbour (pack).
// The entry is published (visible to the scheduler) only once its copies
// and packing have completed.  Caller holds submit_mu, not mu.
int pool_acquire(pmp_server* s, uint64_t id, const uint8_t* pixels, int W, int H, int wrap, bool pack, Request* r) {
    pmp_ctx* ctx = s->ctx;
    const size_t bytes = (size_t)W * H;
    bool fresh = false;
    uint8_t* gray = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = ctx->images.find(id);
        if (it != ctx->images.end()) {
            if (it->second.W != W || it->second.H != H)
                return srv_error(s, PMP_EINVAL, "pooled image " + std::to_string(id) + " has another size");
            it->second.pins++;
            it->second.last_use = ++ctx->pool_clock;
            r->pins.push_back(id);
            if (!pack || it->second.quads) return PMP_OK;
            gray = it->second.gray;   // pinned: stays while we pack it
            int rc = reserve(s, lk, (int64_t)Image::quad_bytes(W, H));
            if (rc) return rc;
        } else {
            if (!pixels)
                return srv_error(s, PMP_ENOTFOUND, "image " + std::to_string(id) + " is not pooled and no pixels given");
            int rc = reserve(s, lk, (int64_t)(bytes + (pack ? Image::quad_bytes(W, H) : 0)));
            if (rc) return rc;
            fresh = true;
        }
    }
    Image im;
    im.W = W;
    im.H = H;
    im.bytes = bytes;
    int rc = PMP_OK;
    if (fresh) {
        hipError_t e = hipMalloc(&im.gray, bytes);
        if (e == hipSuccess) e = hipMemcpyAsync(im.gray, pixels, bytes, hipMemcpyHostToDevice, s->copy);
        if (e != hipSuccess) rc = PMP_EDEVICE;
    } else {
        im.gray = gray;
    }
    if (rc == PMP_OK && pack) {
        std::lock_guard<std::mutex> lk(s->mu);   // ensure_quads may report through the context
        rc = ensure_quads(ctx, im, wrap, s->copy);
    }
    if (rc == PMP_OK && hipStreamSynchronize(s->copy) != hipSuccess) rc = PMP_EDEVICE;
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc != PMP_OK) {
        if (fresh && im.gray) (void)hipFree(im.gray);
        if (im.quads) (void)hipFree(im.quads);
        s->used -= (int64_t)((fresh ? bytes : 0) + (pack ? Image::quad_bytes(W, H) : 0));
        return srv_error(s, PMP_EDEVICE, "upload of image " + std::to_string(id) + " failed");
    }
    if (fresh) {
        im.pins = 1;
        im.last_use = ++ctx->pool_clock;
        ctx->images[id] = im;
        r->pins.push_back(id);
        s->uploads++;
    } else {
        Image& pooled = ctx->images[id];
        pooled.quads = im.quads;
        pooled.quads_wrap = im.quads_wrap;
    }
    return PMP_OK;
}

}  // namespace
}  // namespace pmp

extern "C" {

int pmp_server_create(pmp_server** out, const pmp_server_opts* opts) {
    if (!out) return PMP_EINVAL;
    *out = nullptr;
    pmp_server_opts o;
    memset(&o, 0, sizeof(o));
    if (opts) o = *opts;
    if (o.max_slots == 0) o.max_slots = 16;
    if (o.hbm_budget == 0) o.hbm_budget = 64ll << 30;
    if (o.config.tile_rows == 0) o.config = pmp_batch_config{8, 0, 1, 1};
    if (o.max_slots < 1 || o.max_slots > kMaxSlots || o.hbm_budget < 0 || o.edge_mode < 0 || o.edge_mode > 2 ||
        o.inflight_kpx < 0 ||
        (o.config.tile_rows != 2 && o.config.tile_rows != 4 && o.config.tile_rows != 8))
        return PMP_EINVAL;
    pmp_server* s = new (std::nothrow) pmp_server();
    if (!s) return PMP_ENOMEM;
    s->opts = o;
    s->max_slots = o.max_slots;
    s->budget = o.hbm_budget;
    s->t0 = std::chrono::steady_clock::now();
    int rc = pmp_create(&s->ctx, o.device);
    QueueState* q = nullptr;
    if (!rc) rc = queue_state(s->ctx, q);
    if (!rc) rc = desc_alloc(s->ctx, s->desc[0], kMaxSlots);
    if (!rc) rc = desc_alloc(s->ctx, s->desc[1], kMaxSlots);
    if (!rc && !s->ctx->counters) {
        if (hipMalloc(&s->ctx->counters, 4 * sizeof(unsigned long long)) != hipSuccess) rc = PMP_ENOMEM;
    }
    if (!rc && (hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&s->step_done[0], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&s->step_done[1], hipEventDisableTiming) != hipSuccess))
        rc = PMP_EDEVICE;
    if (rc) {
        if (s->ctx) {
            desc_free(s->desc[0]);
            desc_free(s->desc[1]);
            pmp_destroy(s->ctx);
        }
        delete s;
        return rc;
    }
    s->slot_req.assign(s->max_slots, nullptr);
    s->slot_buf.resize(s->max_slots);
    s->th = std::thread(scheduler, s);
    *out = s;
    return PMP_OK;
}

void pmp_server_destroy(pmp_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->paused = false;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    for (auto& kv : s->reqs) {
        give_host(s, kv.second->out);
        give_host(s, kv.second->lm);
        delete kv.second;
    }
    for (auto& b : s->free_host) (void)hipHostFree(b.p);
    for (auto& b : s->slot_buf) free_slot(b);
    desc_free(s->desc[0]);
    desc_free(s->desc[1]);
    for (hipEvent_t e : s->step_done)
        if (e) (void)hipEventDestroy(e);
    if (s->copy) (void)hipStreamDestroy(s->copy);
    pmp_destroy(s->ctx);
    delete s;
}

const char* pmp_server_last_error(pmp_server* s) { return s ? s->last_error.c_str() : "null server"; }

int pmp_server_submit(pmp_server* s, const pmp_request* req, const pmp_params* params, uint64_t* ticket) {
    if (!s || !req || !params || !ticket) return PMP_EINVAL;
    std::lock_guard<std::mutex> sub(s->submit_mu);
    (void)hipSetDevice(s->ctx->device);
    const pmp_keyframe& kf = req->kf;
    const int W = req->width, H = req->height;
    if (W <= 0 || H <= 0) return srv_error(s, PMP_EINVAL, "width and height must be positive");
    if (kf.n_views < 1 || kf.n_views > kMaxViews) return srv_error(s, PMP_EINVAL, "n_views must be in [1, 10]");
    if (params->edge_mode != s->opts.edge_mode) return srv_error(s, PMP_EINVAL, "edge_mode differs from the server's");
    if (params->iterations > kMaxPasses / 2) return srv_error(s, PMP_EINVAL, "iterations must be <= 64");
    if (req->n_landmarks < 0 || req->n_landmarks > 0x7FFFFFFF || (req->n_landmarks > 0 && !req->landmark_xyz))
        return srv_error(s, PMP_EINVAL, "n_landmarks must be in [0, 2^31) with landmark_xyz given");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (check_params(s->ctx, params) != PMP_OK) return srv_error(s, PMP_EINVAL, s->ctx->last_error);
    }
    if ((int64_t)((size_t)W * H * SlotBuf::kBytesPerPixel) > s->budget)
        return srv_error(s, PMP_ENOMEM, "keyframe maps larger than the HBM budget");
    Request* r = new (std::nothrow) Request();
    if (!r) return PMP_ENOMEM;
    r->kf = kf;
    r->p = *params;
    r->W = W;
    r->H = H;
    r->last_stage = 2 * params->iterations + 1;
    const int wrap = params->edge_mode == PMP_EDGE_WRAP ? 1 : 0;
    int rc = pool_acquire(s, kf.ref_id, req->ref_pixels, W, H, wrap, false, r);
    for (int v = 0; v < kf.n_views && rc == PMP_OK; ++v)
        rc = pool_acquire(s, kf.view_ids[v], req->view_pixels[v], W, H, wrap, true, r);
    if (rc == PMP_OK && kf.mask_id >= 0) rc = pool_acquire(s, (uint64_t)kf.mask_id, req->mask_pixels, W, H, wrap, false, r);
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc == PMP_OK && req->n_landmarks > 0) {
        const size_t bytes = 3 * (size_t)req->n_landmarks * sizeof(double);
        r->lm = take_host(s, bytes);
        if (!r->lm.p) rc = srv_error(s, PMP_ENOMEM, "pinned host landmark buffer");
        else {
            memcpy(r->lm.p, req->landmark_xyz, bytes);
            r->n_lm = req->n_landmarks;
        }
    }
    if (rc != PMP_OK) {
        give_host(s, r->lm);
        unpin(s, r);
        delete r;
        s->cv_room.notify_all();
        return rc;
    }
    r->ticket = s->next_ticket++;
    r->t_submit = s->now_ms();
    s->reqs[r->ticket] = r;
    s->queue.push_back(r);
    s->submitted++;
    *ticket = r->ticket;
    s->cv_work.notify_all();
    return PMP_OK;
}

int pmp_server_wait(pmp_server* s, uint64_t ticket, int32_t timeout_ms) {
    if (!s) return PMP_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return srv_error(s, PMP_ENOTFOUND, "unknown ticket");
    Request* r = it->second;
    auto done = [r] { return r->state == kDone || r->state == kFailed; };
    if (timeout_ms < 0) s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) return PMP_EBUSY;
    return PMP_OK;
}

int pmp_server_fetch(pmp_server* s, uint64_t ticket, float* plane4, float* depth, float* score, int32_t* nghbr,
                     pmp_request_info* info) {
    if (!s) return PMP_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return srv_error(s, PMP_ENOTFOUND, "unknown ticket");
    Request* r = it->second;
    if (r->state != kDone && r->state != kFailed) return srv_error(s, PMP_EBUSY, "request not finished");
    if (info) {
        info->submit_ms = r->t_submit;
        info->admit_ms = r->t_admit;
        info->done_ms = r->t_done;
        info->steps = r->steps;
        info->width = r->W;
        info->height = r->H;
        info->n_views = r->kf.n_views;
    }
    int rc = r->rc;
    if (r->state == kFailed) s->last_error = r->error;
    else {
        const size_t P = (size_t)r->W * r->H;
        const uint8_t* o = r->out.p;
        if (plane4) memcpy(plane4, o, P * 16);
        if (depth) memcpy(depth, o + P * 16, P * 4);
        if (score) memcpy(score, o + P * 20, P * 4);
        if (nghbr) memcpy(nghbr, o + P * 24, P * 4);
    }
    give_host(s, r->out);
    give_host(s, r->lm);
    s->reqs.erase(it);
    delete r;
    return rc;
}

int pmp_server_pause(pmp_server* s, int32_t paused) {
    if (!s) return PMP_EINVAL;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->paused = paused != 0;
    }
    s->cv_work.notify_all();
    return PMP_OK;
}

int pmp_server_stats(pmp_server* s, pmp_server_info* out) {
    if (!s || !out) return PMP_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    memset(out, 0, sizeof(*out));
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->steps = s->steps;
    int64_t pb = 0;
    for (auto& kv : s->ctx->images) pb += (int64_t)image_bytes(kv.second);
    out->pool_images = (int64_t)s->ctx->images.size();
    out->pool_bytes = pb;
    int64_t mb = 0;
    for (auto& b : s->slot_buf) mb += (int64_t)(b.cap * SlotBuf::kBytesPerPixel);
    out->maps_bytes = mb;
    out->uploads = s->uploads;
    out->evictions = s->evictions;
    int fl = 0;
    for (int k = 0; k < s->max_slots; ++k) fl += s->slot_req[k] ? 1 : 0;
    out->in_flight = fl;
    out->queued = (int32_t)s->queue.size();
    out->max_in_flight = s->max_in_flight;
    return PMP_OK;
}

}  // extern "C"
