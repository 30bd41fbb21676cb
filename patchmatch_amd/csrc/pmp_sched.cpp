// pmp_sched.cpp -- the served form's scheduler (see pmp_sched.h).
#include "pmp_sched.h"

#include <string.h>

#include <algorithm>

namespace pmp {
namespace sched {

namespace {
constexpr size_t kOutBytesPerPixel = 28;   // plane 16 | depth 4 | score 4 | nghbr 4
}

Server::Server(Executor* ex, const Options& o) : ex_(ex), o_(o), t0_(std::chrono::steady_clock::now()) {
    slot_req_.assign(o_.max_slots, nullptr);
    slot_cap_.assign(o_.max_slots, 0);
    th_ = std::thread([this] { loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
        paused_ = false;
    }
    cv_work_.notify_all();
    if (th_.joinable()) th_.join();
    for (auto& kv : reqs_) {
        give_host(kv.second->out);
        give_host(kv.second->lm);
        give_host(kv.second->prev);
        delete kv.second;
    }
    for (auto& b : free_host_) ex_->host_free(b.p);
}

double Server::now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
}

bool Server::busy_locked() const {
    for (const Req* r : slot_req_)
        if (r) return true;
    return false;
}

Server::HostBuf Server::take_host(size_t bytes) {
    for (size_t k = 0; k < free_host_.size(); ++k)
        if (free_host_[k].cap >= bytes) {
            HostBuf b = free_host_[k];
            free_host_.erase(free_host_.begin() + (long)k);
            return b;
        }
    HostBuf b;
    b.p = ex_->host_alloc(bytes);
    b.cap = b.p ? bytes : 0;
    return b;
}

void Server::give_host(HostBuf& b) {
    if (b.p) free_host_.push_back(b);
    b = HostBuf();
}

void Server::unpin(Req* r) {
    for (uint64_t id : r->pins) {
        auto it = pool_.find(id);
        if (it != pool_.end() && it->second.pins > 0) it->second.pins--;
    }
    r->pins.clear();
}

// Evicts unpinned pool images, least recently used first, until `need` more
// bytes fit in the budget.  Caller holds mu_.
bool Server::make_room(int64_t need) {
    while (used_ + need > o_.budget) {
        auto victim = pool_.end();
        for (auto it = pool_.begin(); it != pool_.end(); ++it)
            if (it->second.pins == 0 && (victim == pool_.end() || it->second.last_use < victim->second.last_use))
                victim = it;
        if (victim == pool_.end()) return false;
        used_ -= victim->second.bytes;
        ex_->drop(victim->first);   // no queued or running request reads an unpinned image
        pool_.erase(victim);
        evictions_++;
    }
    return true;
}

// Waits until `need` bytes fit in the budget (evicting unpinned images), then
// reserves them.  Caller holds lk on mu_.
int Server::reserve(std::unique_lock<std::mutex>& lk, int64_t need, std::string& err) {
    if (need > o_.budget) {
        err = "image larger than the HBM budget";
        return PMP_ENOMEM;
    }
    while (!make_room(need)) {
        if (!((!queue_.empty() && !paused_) || busy_locked())) {
            err = "HBM budget cannot hold the request's images";
            return PMP_ENOMEM;
        }
        cv_room_.wait(lk);
    }
    used_ += need;
    return PMP_OK;
}

// Records why request r ends (the first reason stands) and retires it once
// no issued step that includes it is still running.  Caller holds mu_.
void Server::end_request(Req* r, int rc, const std::string& msg) {
    if (!r->ending) {
        r->ending = true;
        r->rc = rc;
        r->error = msg;
    }
    if (r->inflight == 0) retire(r);
}

// Frees what request r holds on the device side (slot, pins, landmark
// buffer) and makes its end visible.  Caller holds mu_; r->inflight == 0.
void Server::retire(Req* r) {
    if (r->slot >= 0) {
        slot_req_[r->slot] = nullptr;
        r->slot = -1;
    }
    unpin(r);
    give_host(r->lm);
    r->job.lm = nullptr;
    give_host(r->prev);
    r->job.prev_plane4 = nullptr;
    r->job.prev_depth = nullptr;
    r->t_done = now_ms();
    if (r->rc == PMP_OK) {
        r->state = kDone;
        completed_++;
    } else if (r->rc == PMP_ECANCELED) {
        r->state = kCancelled;
        cancelled_++;
    } else {
        r->state = kFailed;
        failed_++;
    }
    cv_done_.notify_all();
    cv_room_.notify_all();
}

// Admission of queued requests into free slots, oldest first, within the
// in-flight pixel budget.  Caller holds mu_.
void Server::admit() {
    while (!queue_.empty() && !paused_) {
        Req* r = queue_.front();
        const int64_t P = (int64_t)r->job.W * r->job.H;
        // best fit among the free slots: the smallest whose maps already hold
        // P pixels, else the largest (least growth), so that one large request
        // does not grow every slot's maps in turn
        int slot = -1;
        int64_t inflight = 0;
        for (int k = 0; k < o_.max_slots; ++k) {
            if (slot_req_[k]) {
                inflight += (int64_t)slot_req_[k]->job.W * slot_req_[k]->job.H;
                continue;
            }
            if (slot < 0) { slot = k; continue; }
            const int64_t c = slot_cap_[k], b = slot_cap_[slot];
            if (c >= P ? (b < P || c < b) : (b < P && c > b)) slot = k;
        }
        if (slot < 0) return;
        if (o_.inflight_kpx > 0 && inflight > 0 && inflight + P > 1000ll * o_.inflight_kpx) return;
        // tail admission: without a backlog (one request queued), a request
        // joins the ones in flight only for their last tail_stages stages, so
        // that they finish about when they would alone (sharing every pass
        // launch would delay the oldest); with a backlog, admission is free
        if (o_.tail_stages > 0 && inflight > 0 && queue_.size() < 2) {
            for (int k = 0; k < o_.max_slots; ++k) {
                const Req* q = slot_req_[k];
                if (q && !q->ending && q->job.last_stage + 1 - q->stage > o_.tail_stages) return;
            }
        }
        if (P > slot_cap_[slot]) {
            const int64_t grow = ex_->map_bytes(P) - ex_->map_bytes(slot_cap_[slot]);
            if (!make_room(grow)) {
                if (busy_locked()) return;   // wait for requests in flight to finish
                // nothing will free memory: the pinned images of queued requests leave no room
                queue_.pop_front();
                end_request(r, PMP_ENOMEM, "HBM budget cannot hold the keyframe maps next to the queued requests' images");
                continue;
            }
            std::string err;
            used_ -= ex_->map_bytes(slot_cap_[slot]);
            slot_cap_[slot] = 0;
            if (ex_->slot_grow(slot, P, err) != PMP_OK) {
                queue_.pop_front();
                end_request(r, PMP_ENOMEM, "keyframe maps: " + err);
                continue;
            }
            slot_cap_[slot] = P;
            used_ += ex_->map_bytes(P);
        }
        const size_t out_bytes = (size_t)P * kOutBytesPerPixel;
        if (!r->out.p || r->out.cap < out_bytes) {
            give_host(r->out);
            r->out = take_host(out_bytes);
        }
        queue_.pop_front();
        if (!r->out.p) {
            end_request(r, PMP_ENOMEM, "pinned host result buffer");
            continue;
        }
        r->job.out = r->out.p;
        r->state = kActive;
        r->slot = slot;
        r->stage = 0;
        r->t_admit = now_ms();
        slot_req_[slot] = r;
    }
}

// Builds and issues one step into descriptor set w.  Returns false when no
// request has a stage to run.
bool Server::issue_step(int w) {
    std::vector<StepItem> items;
    std::vector<Req*> owners;
    {
        std::lock_guard<std::mutex> lk(mu_);
        admit();
        int active = 0;
        for (int k = 0; k < o_.max_slots; ++k) {
            Req* r = slot_req_[k];
            if (!r) continue;
            ++active;
            // ending, or its finalize is in the step before this one
            if (r->ending || r->stage > r->job.last_stage) continue;
            items.push_back(StepItem{k, r->stage, &r->job});
            owners.push_back(r);
            r->inflight++;   // keeps the request (and its job) alive through step_issue
            r->steps++;
            r->stage++;
        }
        if (active > max_in_flight_) max_in_flight_ = active;
        if (items.empty()) return false;
        steps_++;
    }
    std::vector<int> irc(items.size(), PMP_OK);
    std::vector<std::string> ierr(items.size());
    std::string err;
    const int rc = ex_->step_issue(w, items, irc, ierr, err);
    std::lock_guard<std::mutex> lk(mu_);
    if (rc != PMP_OK) {
        // nothing of this step runs, and the device stream is in doubt: every
        // request in flight fails (those in the other pending step when it is
        // collected)
        for (Req* r : owners) r->inflight--;
        for (int k = 0; k < o_.max_slots; ++k)
            if (slot_req_[k]) end_request(slot_req_[k], rc, err);
        return true;
    }
    std::vector<StepItem> issued;
    for (size_t k = 0; k < items.size(); ++k) {
        if (irc[k] != PMP_OK) {
            owners[k]->inflight--;
            end_request(owners[k], irc[k], ierr[k]);
        } else {
            issued.push_back(items[k]);
        }
    }
    step_items_[w] = issued;
    step_pending_[w] = !issued.empty();
    return true;
}

// Completions of the step issued into descriptor set w.
void Server::collect(int w) {
    if (!step_pending_[w]) return;
    std::string err;
    const int rc = ex_->step_wait(w, err);
    std::lock_guard<std::mutex> lk(mu_);
    step_pending_[w] = false;
    for (const StepItem& it : step_items_[w]) {
        Req* r = slot_req_[it.slot];   // still there: its inflight count kept it
        r->inflight--;
        if (rc != PMP_OK) end_request(r, rc, err);
        else if (it.stage == r->job.last_stage) end_request(r, PMP_OK, std::string());
        else if (r->ending && r->inflight == 0) retire(r);
    }
    step_items_[w].clear();
    cv_room_.notify_all();
}

void Server::loop() {
    ex_->bind_thread();
    int w = 0;
    for (;;) {
        const bool issued = issue_step(w);
        // collect the previous step while this one runs
        collect(w ^ 1);
        if (!issued) {
            collect(w);
            std::unique_lock<std::mutex> lk(mu_);
            if (busy_locked()) continue;
            if (stop_ && (queue_.empty() || paused_)) break;
            cv_work_.wait(lk, [this] { return stop_ || (!queue_.empty() && !paused_); });
            continue;
        }
        w ^= 1;
    }
    collect(0);
    collect(1);
}

// Pins the pool entry of `id` for request r, uploading it from host pixels
// if absent and packing its neighbour copy if it serves as a neighbour.  The
// entry becomes visible to the scheduler only once its upload and packing
// have completed.  Caller holds submit_mu_, not mu_.
int Server::acquire(uint64_t id, const uint8_t* pixels, int W, int H, bool pack, Req* r, std::string& err) {
    bool fresh = false, need_pack = false;
    int64_t reserved = 0;
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = pool_.find(id);
        if (it != pool_.end()) {
            if (it->second.W != W || it->second.H != H) {
                err = "pooled image " + std::to_string(id) + " has another size";
                return PMP_EINVAL;
            }
            it->second.pins++;
            it->second.last_use = ++clock_;
            r->pins.push_back(id);
            if (!pack || it->second.packed) return PMP_OK;
            reserved = ex_->image_bytes(W, H, true) - ex_->image_bytes(W, H, false);
            const int rc = reserve(lk, reserved, err);
            if (rc) return rc;
            need_pack = true;
        } else {
            if (!pixels) {
                err = "image " + std::to_string(id) + " is not pooled and no pixels given";
                return PMP_ENOTFOUND;
            }
            reserved = ex_->image_bytes(W, H, pack);
            const int rc = reserve(lk, reserved, err);
            if (rc) return rc;
            fresh = true;
            need_pack = pack;
        }
    }
    int rc = PMP_OK;
    if (fresh) rc = ex_->upload(id, pixels, W, H, err);
    if (rc == PMP_OK && need_pack) rc = ex_->pack(id, o_.edge_mode == PMP_EDGE_WRAP ? 1 : 0, err);
    std::lock_guard<std::mutex> lk(mu_);
    if (rc != PMP_OK) {
        if (fresh) ex_->drop(id);
        used_ -= reserved;
        cv_room_.notify_all();
        return rc;
    }
    if (fresh) {
        Entry e;
        e.W = W;
        e.H = H;
        e.packed = pack;
        e.pins = 1;
        e.last_use = ++clock_;
        e.bytes = reserved;
        pool_[id] = e;
        r->pins.push_back(id);
        uploads_++;
    } else {
        Entry& e = pool_[id];
        e.packed = true;
        e.bytes += reserved;
    }
    return PMP_OK;
}

int Server::submit(const pmp_request& req, const pmp_params& p, uint64_t* ticket, std::string& err) {
    std::lock_guard<std::mutex> sub(submit_mu_);
    const pmp_keyframe& kf = req.kf;
    const int W = req.width, H = req.height;
    if (W <= 0 || H <= 0) {
        err = "width and height must be positive";
        return PMP_EINVAL;
    }
    if (kf.n_views < 1 || kf.n_views > PMP_MAX_VIEWS) {
        err = "n_views must be in [1, 10]";
        return PMP_EINVAL;
    }
    if (p.edge_mode != o_.edge_mode) {
        err = "edge_mode differs from the server's";
        return PMP_EINVAL;
    }
    if (req.n_landmarks < 0 || req.n_landmarks > 0x7FFFFFFF || (req.n_landmarks > 0 && !req.landmark_xyz)) {
        err = "n_landmarks must be in [0, 2^31) with landmark_xyz given";
        return PMP_EINVAL;
    }
    const bool warm = req.prev_plane4 || req.prev_depth || req.prev_pose;
    if (warm && !(req.prev_plane4 && req.prev_depth && req.prev_pose)) {
        err = "a warm start needs prev_plane4, prev_depth and prev_pose";
        return PMP_EINVAL;
    }
    if (ex_->check_params(p, err) != PMP_OK) return PMP_EINVAL;
    if (ex_->map_bytes((int64_t)W * H) > o_.budget) {
        err = "keyframe maps larger than the HBM budget";
        return PMP_ENOMEM;
    }
    Req* r = new (std::nothrow) Req();
    if (!r) return PMP_ENOMEM;
    r->job.kf = kf;
    r->job.p = p;
    r->job.W = W;
    r->job.H = H;
    r->job.last_stage = 2 * p.iterations + 1;
    int rc = acquire(kf.ref_id, req.ref_pixels, W, H, false, r, err);
    for (int v = 0; v < kf.n_views && rc == PMP_OK; ++v)
        rc = acquire(kf.view_ids[v], req.view_pixels[v], W, H, true, r, err);
    if (rc == PMP_OK && kf.mask_id >= 0) rc = acquire((uint64_t)kf.mask_id, req.mask_pixels, W, H, false, r, err);
    if (rc == PMP_OK && warm) {
        // the previous result into pinned memory (copied without the lock:
        // r is not visible to the scheduler before it is queued)
        const size_t P = (size_t)W * H;
        {
            std::lock_guard<std::mutex> l2(mu_);
            r->prev = take_host(20 * P);
        }
        if (!r->prev.p) {
            err = "pinned host buffer of the previous result";
            rc = PMP_ENOMEM;
        } else {
            memcpy(r->prev.p, req.prev_plane4, 16 * P);
            memcpy(r->prev.p + 16 * P, req.prev_depth, 4 * P);
            r->job.prev_plane4 = (const float*)r->prev.p;
            r->job.prev_depth = (const float*)(r->prev.p + 16 * P);
            memcpy(r->job.prev_pose, req.prev_pose->R, 9 * sizeof(double));
            memcpy(r->job.prev_pose + 9, req.prev_pose->t, 3 * sizeof(double));
        }
    }
    std::lock_guard<std::mutex> lk(mu_);
    if (rc == PMP_OK && req.n_landmarks > 0) {
        const size_t bytes = 3 * (size_t)req.n_landmarks * sizeof(double);
        r->lm = take_host(bytes);
        if (!r->lm.p) {
            err = "pinned host landmark buffer";
            rc = PMP_ENOMEM;
        } else {
            memcpy(r->lm.p, req.landmark_xyz, bytes);
            r->job.lm = (const double*)r->lm.p;
            r->job.n_lm = req.n_landmarks;
        }
    }
    if (rc != PMP_OK) {
        give_host(r->lm);
        give_host(r->prev);
        unpin(r);
        delete r;
        cv_room_.notify_all();
        return rc;
    }
    r->job.ticket = next_ticket_++;
    r->t_submit = now_ms();
    reqs_[r->job.ticket] = r;
    queue_.push_back(r);
    submitted_++;
    *ticket = r->job.ticket;
    cv_work_.notify_all();
    return PMP_OK;
}

int Server::wait(uint64_t ticket, int timeout_ms, std::string& err) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) {
        err = "unknown ticket";
        return PMP_ENOTFOUND;
    }
    Req* r = it->second;
    auto done = [r] { return r->state == kDone || r->state == kFailed || r->state == kCancelled; };
    if (timeout_ms < 0) cv_done_.wait(lk, done);
    else if (!cv_done_.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) return PMP_EBUSY;
    return PMP_OK;
}

int Server::fetch(uint64_t ticket, float* plane4, float* depth, float* score, int32_t* nghbr, pmp_request_info* info,
                  std::string& err) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) {
        err = "unknown ticket";
        return PMP_ENOTFOUND;
    }
    Req* r = it->second;
    if (r->state != kDone && r->state != kFailed && r->state != kCancelled) {
        err = "request not finished";
        return PMP_EBUSY;
    }
    if (info) {
        info->submit_ms = r->t_submit;
        info->admit_ms = r->t_admit;
        info->done_ms = r->t_done;
        info->steps = r->steps;
        info->width = r->job.W;
        info->height = r->job.H;
        info->n_views = r->job.kf.n_views;
    }
    const int rc = r->rc;
    if (r->state != kDone) {
        err = r->error;
    } else {
        const size_t P = (size_t)r->job.W * r->job.H;
        const uint8_t* o = r->out.p;
        if (plane4) memcpy(plane4, o, P * 16);
        if (depth) memcpy(depth, o + P * 16, P * 4);
        if (score) memcpy(score, o + P * 20, P * 4);
        if (nghbr) memcpy(nghbr, o + P * 24, P * 4);
    }
    give_host(r->out);
    give_host(r->lm);
    give_host(r->prev);
    reqs_.erase(it);
    delete r;
    return rc;
}

int Server::cancel(uint64_t ticket, std::string& err) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) {
        err = "unknown ticket";
        return PMP_ENOTFOUND;
    }
    Req* r = it->second;
    if (r->state == kQueued) {
        queue_.erase(std::find(queue_.begin(), queue_.end(), r));
        end_request(r, PMP_ECANCELED, "cancelled");
    } else if (r->state == kActive) {
        end_request(r, PMP_ECANCELED, "cancelled");   // retires once its issued steps have run
    }
    return PMP_OK;
}

int Server::reset(std::string& err) {
    (void)err;
    std::lock_guard<std::mutex> sub(submit_mu_);   // no upload in progress, none starts
    std::unique_lock<std::mutex> lk(mu_);
    while (!queue_.empty()) {
        Req* r = queue_.front();
        queue_.pop_front();
        end_request(r, PMP_ECANCELED, "cancelled by reset");
    }
    for (int k = 0; k < o_.max_slots; ++k)
        if (slot_req_[k]) end_request(slot_req_[k], PMP_ECANCELED, "cancelled by reset");
    cv_work_.notify_all();
    cv_done_.wait(lk, [this] { return !busy_locked(); });
    // no request holds a pin now: the pool empties (the reference clears its
    // keyframe queue and its depth-map database)
    for (auto it = pool_.begin(); it != pool_.end();) {
        if (it->second.pins == 0) {
            used_ -= it->second.bytes;
            ex_->drop(it->first);
            it = pool_.erase(it);
        } else {
            ++it;
        }
    }
    cv_room_.notify_all();
    return PMP_OK;
}

int Server::pause(bool paused) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        paused_ = paused;
    }
    cv_work_.notify_all();
    cv_room_.notify_all();
    return PMP_OK;
}

int Server::stats(pmp_server_info* out) {
    std::lock_guard<std::mutex> lk(mu_);
    memset(out, 0, sizeof(*out));
    out->submitted = submitted_;
    out->completed = completed_;
    out->failed = failed_;
    out->steps = steps_;
    int64_t pb = 0;
    for (auto& kv : pool_) pb += kv.second.bytes;
    out->pool_images = (int64_t)pool_.size();
    out->pool_bytes = pb;
    int64_t mb = 0;
    for (int64_t c : slot_cap_) mb += ex_->map_bytes(c);
    out->maps_bytes = mb;
    out->uploads = uploads_;
    out->evictions = evictions_;
    int fl = 0;
    for (const Req* r : slot_req_) fl += r ? 1 : 0;
    out->in_flight = fl;
    out->queued = (int32_t)queue_.size();
    out->max_in_flight = max_in_flight_;
    out->cancelled = (int32_t)cancelled_;
    return PMP_OK;
}

}  // namespace sched
}  // namespace pmp
