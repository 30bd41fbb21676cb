// pmp_server.hip -- the served form on the GPU: the scheduler of
// pmp_sched.cpp driving the work-queue kernels of pmp_queue.hip.
//
// The executor owns a device context (image pool, slot maps, two descriptor
// sets), a copy stream for uploads and packing, and an event per descriptor
// set.  A step fills the descriptors of its slots, then launches on the
// context stream, each over every slot of its list: the init kernel (plus
// landmark seeding per seeded slot), one red-black pass, the finalize
// kernels, and the copies of finished maps to pinned host memory; an event
// marks its end.  Uploads and packing run on the copy stream from the
// submitting thread and overlap the compute stream.  Slot maps are
// independent, so every request's maps are those of pmp_densify with the same
// params.
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "pmp_internal.h"
#include "pmp_trace.h"
#include "pmp_queue.h"
#include "pmp_sched.h"

namespace pmp {
namespace {

class GpuExec final : public sched::Executor {
  public:
    pmp_ctx* ctx = nullptr;
    hipStream_t copy = nullptr;
    DescSet desc[2];
    hipEvent_t step_done[2] = {nullptr, nullptr};
    std::vector<SlotBuf> slots;
    pmp_batch_config cfg = kDefaultConfig;
    int edge_mode = PMP_EDGE_WRAP;
    // ctx->images and ctx->last_error: touched by the submitting thread
    // (upload, pack) and the scheduler thread (drop, slot_grow, step_issue).
    // The scheduler thread keeps the context's device current; calls on the
    // caller's threads (submit, reset) hold a DeviceScope.
    std::mutex pool_mu;
    // pinned staging of uploads (one copy from pinned memory instead of a
    // pageable one), shared by the submitting threads
    std::mutex stage_mu;
    uint8_t* stage = nullptr;
    size_t stage_cap = 0;

    int init(const pmp_server_opts& o) {
        cfg = o.config;
        edge_mode = o.edge_mode;
        int rc = pmp_create(&ctx, o.device);
        if (rc) return rc;
        DeviceScope dev(ctx->device);
        QueueState* q = nullptr;
        rc = queue_state(ctx, q);
        if (!rc) rc = desc_alloc(ctx, desc[0], kMaxSlots);
        if (!rc) rc = desc_alloc(ctx, desc[1], kMaxSlots);
        if (!rc && !ctx->counters) {
            if (hipMalloc(&ctx->counters, 4 * sizeof(unsigned long long)) != hipSuccess) rc = PMP_ENOMEM;
        }
        if (!rc && (hipStreamCreateWithFlags(&copy, hipStreamNonBlocking) != hipSuccess ||
                    hipEventCreateWithFlags(&step_done[0], hipEventDisableTiming) != hipSuccess ||
                    hipEventCreateWithFlags(&step_done[1], hipEventDisableTiming) != hipSuccess))
            rc = PMP_EDEVICE;
        slots.resize(o.max_slots);
        return rc;
    }

    ~GpuExec() override {
        if (!ctx) return;
        DeviceScope dev(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (auto& b : slots) free_slot(b);
        desc_free(desc[0]);
        desc_free(desc[1]);
        for (hipEvent_t e : step_done)
            if (e) (void)hipEventDestroy(e);
        if (copy) (void)hipStreamDestroy(copy);
        if (stage) (void)hipHostFree(stage);
        pmp_destroy(ctx);   // frees the pooled images
    }

    void bind_thread() override { (void)hipSetDevice(ctx->device); }

    int check_params(const pmp_params& p, std::string& err) override {
        if (p.iterations > kMaxPasses / 2) {
            err = "iterations must be <= 64";
            return PMP_EINVAL;
        }
        std::lock_guard<std::mutex> lk(pool_mu);
        const int rc = pmp::check_params(ctx, &p);
        if (rc) err = ctx->last_error;
        return rc;
    }

    int64_t image_bytes(int W, int H, bool packed) const override {
        return (int64_t)W * H + (packed ? (int64_t)Image::quad_bytes(W, H) : 0);
    }
    int64_t map_bytes(int64_t pixels) const override { return pixels * (int64_t)SlotBuf::kBytesPerPixel; }

    int upload(uint64_t id, const uint8_t* pixels, int W, int H, std::string& err) override {
        TraceRange trace_range("pmp_server upload");
        Image im;
        im.W = W;
        im.H = H;
        im.bytes = (size_t)W * H;
        hipError_t e = hipMalloc(&im.gray, im.bytes);
        if (e == hipSuccess) {
            std::lock_guard<std::mutex> lk(stage_mu);
            if (im.bytes > stage_cap) {
                if (stage) (void)hipHostFree(stage);
                stage = nullptr;
                stage_cap = 0;
                e = hipHostMalloc((void**)&stage, im.bytes, hipHostMallocDefault);
                if (e == hipSuccess) stage_cap = im.bytes;
            }
            if (e == hipSuccess) {
                memcpy(stage, pixels, im.bytes);
                e = hipMemcpyAsync(im.gray, stage, im.bytes, hipMemcpyHostToDevice, copy);
            }
            // the staging buffer is reused by the next upload
            if (e == hipSuccess) e = hipStreamSynchronize(copy);
        }
        if (e != hipSuccess) {
            if (im.gray) (void)hipFree(im.gray);
            err = "upload of image " + std::to_string(id) + ": " + hipGetErrorString(e);
            return PMP_EDEVICE;
        }
        std::lock_guard<std::mutex> lk(pool_mu);
        ctx->images[id] = im;
        return PMP_OK;
    }

    int pack(uint64_t id, int wrap, std::string& err) override {
        std::lock_guard<std::mutex> lk(pool_mu);   // ensure_quads reports through the context
        auto it = ctx->images.find(id);
        if (it == ctx->images.end()) {
            err = "packing an image that is not pooled";
            return PMP_ENOTFOUND;
        }
        int rc = ensure_quads(ctx, it->second, wrap, copy);
        if (rc == PMP_OK && hipStreamSynchronize(copy) != hipSuccess) rc = PMP_EDEVICE;
        if (rc) err = "packing of image " + std::to_string(id) + ": " + ctx->last_error;
        return rc;
    }

    void drop(uint64_t id) override {
        std::lock_guard<std::mutex> lk(pool_mu);
        auto it = ctx->images.find(id);
        if (it == ctx->images.end()) return;
        if (it->second.gray) (void)hipFree(it->second.gray);
        if (it->second.quads) (void)hipFree(it->second.quads);
        ctx->images.erase(it);
    }

    int slot_grow(int k, int64_t pixels, std::string& err) override {
        std::lock_guard<std::mutex> lk(pool_mu);
        SlotBuf& b = slots[k];
        free_slot(b);
        // a W x H request fits any slot of >= W*H pixels: the maps are indexed by the request's own sizes
        const int rc = ensure_slot(ctx, b, (int)pixels, 1);
        if (rc) err = ctx->last_error;
        return rc;
    }

    uint8_t* host_alloc(size_t bytes) override {
        uint8_t* p = nullptr;
        if (hipHostMalloc((void**)&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
        return p;
    }
    void host_free(uint8_t* p) override { (void)hipHostFree(p); }

    int step_issue(int w, const std::vector<sched::StepItem>& items, std::vector<int>& item_rc,
                   std::vector<std::string>& item_err, std::string& err) override {
        TraceRange trace_range("pmp_server step");
        DescSet& d = desc[w];
        int items_init[kMaxSlots] = {0}, items_prop[kMaxSlots] = {0}, items_rows[kMaxSlots] = {0};
        bool post = false;
        std::vector<const sched::StepItem*> fin, seeds, warms;
        int rc = PMP_OK;
        // the descriptors read the pool, and errors are reported through the context
        std::lock_guard<std::mutex> lk(pool_mu);
        {
            if (desc_begin(ctx, d) != PMP_OK) {
                err = ctx->last_error;
                return PMP_EDEVICE;
            }
            for (size_t k = 0; k < items.size(); ++k) {
                const sched::StepItem& it = items[k];
                const sched::Job* j = it.job;
                SlotBuf& b = slots[it.slot];
                b.W = j->W;
                b.H = j->H;
                int its[3];
                const int pass0 = it.stage >= 1 && it.stage < j->last_stage ? it.stage - 1 : 0;
                const int frc = slot_fill(ctx, &j->p, j->kf, b, nullptr, pass0, cfg.tile_rows, ctx->stream,
                                          d.h_slots[it.slot], its);
                if (frc) {
                    item_rc[k] = frc;
                    item_err[k] = ctx->last_error;
                    continue;
                }
                if (it.stage == 0) {
                    items_init[it.slot] = its[0];
                    if (j->n_lm > 0) seeds.push_back(&it);
                    if (j->prev_plane4) warms.push_back(&it);
                } else if (it.stage < j->last_stage) {
                    items_prop[it.slot] = its[1];
                } else {
                    items_rows[it.slot] = its[2];
                    post = post || j->p.postprocess;
                    fin.push_back(&it);
                }
            }
        }
        const int n = (int)slots.size();
        fill_list(d.h_hdr->init, n, items_init);
        fill_list(d.h_hdr->prop, n, items_prop);
        fill_list(d.h_hdr->rows, n, items_rows);
        QueueRun run;
        rc = queue_run(ctx, *ctx->queue, cfg, edge_mode == PMP_EDGE_ZERO, post, run);
        if (!rc) rc = desc_upload(ctx, d, n, ctx->stream);
        if (!rc) rc = q_clear(ctx, ctx->stream, d, 1);
        if (!rc && d.h_hdr->init.n) rc = q_init(ctx, ctx->stream, d, run);
        for (const sched::StepItem* it : seeds) {
            if (rc) break;
            const sched::Job* j = it->job;
            rc = preinit_landmarks_slot(ctx, d.d_slots + it->slot, j->W, j->H, (const double*)&j->kf.ref_pose,
                                        j->n_lm, j->lm, edge_mode == PMP_EDGE_ZERO);
        }
        // then the warm starts (PreInitResultWarp after PreInitSparsePCL2,
        // depthmap.cc:669-719)
        for (const sched::StepItem* it : warms) {
            if (rc) break;
            const sched::Job* j = it->job;
            rc = preinit_warp_slot(ctx, d.d_slots + it->slot, j->W, j->H, (const double*)&j->kf.ref_pose,
                                   j->prev_pose, j->prev_plane4, j->prev_depth, edge_mode == PMP_EDGE_ZERO);
        }
        if (!rc && d.h_hdr->prop.n) rc = q_prop(ctx, ctx->stream, d, run, 0);
        if (!rc && d.h_hdr->rows.n) rc = q_finalize(ctx, ctx->stream, d, run);
        for (const sched::StepItem* it : fin) {
            if (rc) break;
            const SlotBuf& b = slots[it->slot];
            const size_t P = (size_t)it->job->W * it->job->H;
            uint8_t* o = it->job->out;
            hipError_t e = hipMemcpyAsync(o, b.oplane, P * 16, hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(o + P * 16, b.odepth, P * 4, hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(o + P * 20, b.oscore, P * 4, hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(o + P * 24, b.onghbr, P * 4, hipMemcpyDeviceToHost, ctx->stream);
            if (e != hipSuccess) rc = check_hip(ctx, e, "result copy");
        }
        if (!rc && hipEventRecord(step_done[w], ctx->stream) != hipSuccess) rc = PMP_EDEVICE;
        if (rc) {
            (void)hipStreamSynchronize(ctx->stream);
            err = ctx->last_error;
        }
        return rc;
    }

    int step_wait(int w, std::string& err) override {
        const hipError_t e = hipEventSynchronize(step_done[w]);
        if (e != hipSuccess) {
            err = hipGetErrorString(e);
            return PMP_EDEVICE;
        }
        return PMP_OK;
    }
};

// The reason of the last failing pmp_server_* call of this thread.
thread_local std::string t_server_error;

int fail(int rc, const std::string& msg) {
    t_server_error = msg;
    return rc;
}

}  // namespace
}  // namespace pmp

using namespace pmp;

struct pmp_server {
    GpuExec exec;
    sched::Server* srv = nullptr;
};

extern "C" {

int pmp_server_create(pmp_server** out, const pmp_server_opts* opts) {
    if (!out) return PMP_EINVAL;
    *out = nullptr;
    pmp_server_opts o;
    memset(&o, 0, sizeof(o));
    if (opts) o = *opts;
    if (o.max_slots == 0) o.max_slots = 16;
    if (o.hbm_budget == 0) o.hbm_budget = 64ll << 30;
    if (o.config.tile_rows == 0) o.config = kDefaultConfig;
    if (o.max_slots < 1 || o.max_slots > kMaxSlots || o.hbm_budget < 0 || o.edge_mode < 0 || o.edge_mode > 2 ||
        o.inflight_kpx < 0 || o.tail_stages < 0 ||
        (o.config.tile_rows != 2 && o.config.tile_rows != 4 && o.config.tile_rows != 8))
        return fail(PMP_EINVAL, "invalid server options");
    pmp_server* s = new (std::nothrow) pmp_server();
    if (!s) return PMP_ENOMEM;
    const int rc = s->exec.init(o);
    if (rc) {
        delete s;
        return fail(rc, "device context of the server");
    }
    sched::Options so;
    so.max_slots = o.max_slots;
    so.budget = o.hbm_budget;
    so.edge_mode = o.edge_mode;
    so.inflight_kpx = o.inflight_kpx;
    so.tail_stages = o.tail_stages;
    s->srv = new (std::nothrow) sched::Server(&s->exec, so);
    if (!s->srv) {
        delete s;
        return PMP_ENOMEM;
    }
    *out = s;
    return PMP_OK;
}

void pmp_server_destroy(pmp_server* s) {
    if (!s) return;
    delete s->srv;   // finishes the queued requests and joins the scheduler
    delete s;        // then the executor releases the device
}

const char* pmp_server_last_error(pmp_server* s) {
    (void)s;
    return t_server_error.c_str();
}

int pmp_server_submit(pmp_server* s, const pmp_request* req, const pmp_params* params, uint64_t* ticket) {
    if (!s || !req || !params || !ticket) return PMP_EINVAL;
    DeviceScope dev(s->exec.ctx->device);   // uploads and packing run on this thread
    std::string err;
    const int rc = s->srv->submit(*req, *params, ticket, err);
    return rc ? fail(rc, err) : PMP_OK;
}

int pmp_server_wait(pmp_server* s, uint64_t ticket, int32_t timeout_ms) {
    if (!s) return PMP_EINVAL;
    std::string err;
    const int rc = s->srv->wait(ticket, timeout_ms, err);
    return (rc && rc != PMP_EBUSY) ? fail(rc, err) : rc;
}

int pmp_server_fetch(pmp_server* s, uint64_t ticket, float* plane4, float* depth, float* score, int32_t* nghbr,
                     pmp_request_info* info) {
    if (!s) return PMP_EINVAL;
    std::string err;
    const int rc = s->srv->fetch(ticket, plane4, depth, score, nghbr, info, err);
    return rc ? fail(rc, err) : PMP_OK;
}

int pmp_server_cancel(pmp_server* s, uint64_t ticket) {
    if (!s) return PMP_EINVAL;
    std::string err;
    const int rc = s->srv->cancel(ticket, err);
    return rc ? fail(rc, err) : PMP_OK;
}

int pmp_server_reset(pmp_server* s) {
    if (!s) return PMP_EINVAL;
    DeviceScope dev(s->exec.ctx->device);   // the pool's images are freed on this thread
    std::string err;
    const int rc = s->srv->reset(err);
    return rc ? fail(rc, err) : PMP_OK;
}

int pmp_server_pause(pmp_server* s, int32_t paused) {
    if (!s) return PMP_EINVAL;
    return s->srv->pause(paused != 0);
}

int pmp_server_stats(pmp_server* s, pmp_server_info* out) {
    if (!s || !out) return PMP_EINVAL;
    return s->srv->stats(out);
}

}  // extern "C"
