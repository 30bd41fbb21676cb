// pmp_estimator.hip -- the one-keyframe API (DepthmapEstimator,
// depthmap.cc:541-806): keyframe selection, kernel arguments, and the stage
// calls, each launching the work-queue kernels of pmp_queue.hip over a
// descriptor set with one slot (the context workspace):
//   pmp_random_initialization  k_init_q       (PatchMatchCuda.cu:210-270)
//   pmp_score_plane_depth      k_score_q      (PatchMatchCuda.cu:255-270)
//   pmp_red_black_pass         k_propagate_q  one launch per red-black pass
//                                             (PatchMatchCuda.cu:273-407)
//   pmp_finalize               k_copyback_q, k_median_q x2, k_postfilter_q
//                                             (depthmap.cc:673-687, :731-760,
//                                             :1169-1208)
//   pmp_densify                all of the above, replayed from a hipGraph
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdio>
#include <string>

#include "pmp_trace.h"
#include "pmp_internal.h"
#include "pmp_math.h"
#include "pmp_queue.h"

namespace pmp {

// ---------------------------------------------------------------- host helpers
int check_params(pmp_ctx* ctx, const pmp_params* p) {
    if (p->patch_half < 3 || p->patch_stride < 1 || p->patch_half > kMaxHalf)
        return set_error(ctx, PMP_EINVAL, "patch_half must be in [3, 10] (propagation reads 3 pixels away)");
    if (p->n_refine < 0 || p->n_refine > kMaxRefine) return set_error(ctx, PMP_EINVAL, "n_refine must be in [0, 8]");
    if (p->agg_mode < 0 || p->agg_mode > 3 || p->edge_mode < 0 || p->edge_mode > 2 || p->coverage_mode < 0 ||
        p->coverage_mode > 1 || p->init_mode < 0 || p->init_mode > 1)
        return set_error(ctx, PMP_EINVAL, "unknown mode value");
    if (!(p->sigma_color > 0.0f) || !(p->sigma_space > 0.0f))
        return set_error(ctx, PMP_EINVAL, "sigma_color and sigma_space must be positive");
    if (p->iterations < 0) return set_error(ctx, PMP_EINVAL, "iterations must be >= 0");
    return PMP_OK;
}

// Kernel arguments of one keyframe: reference image ref_id, N neighbours
// (pooled ids), world->camera poses as 12 doubles (R row-major, t).  Packs the
// neighbours' 2x2 words on `stream` where the pool does not hold them yet.
int build_args(pmp_ctx* ctx, const pmp_params* p, uint64_t ref_id, int N, const uint64_t* view_ids,
               const double* ref_pose, const double* view_poses, hipStream_t stream, KfArgs& a) {
    int rc = check_params(ctx, p);
    if (rc) return rc;
    if (N < 1 || N > kMaxViews) return set_error(ctx, PMP_EINVAL, "n_views must be in [1, 10]");
    auto ir = ctx->images.find(ref_id);
    if (ir == ctx->images.end()) return set_error(ctx, PMP_ENOTFOUND, "unknown reference image id");
    memset(&a, 0, sizeof(a));
    a.W = ir->second.W; a.H = ir->second.H; a.N = N;
    a.half = p->patch_half; a.stride = p->patch_stride;
    int T = 0;
    for (int k1 = -a.half; k1 <= a.half; k1 += a.stride)
        for (int k2 = -a.half; k2 <= a.half; k2 += a.stride) {
            if (T >= kMaxTaps) return set_error(ctx, PMP_EINVAL, "too many taps (max 121)");
            a.Ws[T++] = (float)exp(-(double)(k1 * k1 + k2 * k2) /
                                   (2.0 * (double)p->sigma_space * (double)p->sigma_space));
        }
    a.T = T;
    {
        const int n = 2 * a.half / a.stride + 1;   // taps per row (and per column)
        a.tc = (n / 2) * n + n / 2;
        a.kc = -a.half + (n / 2) * a.stride;
        a.klast = -a.half + (n - 1) * a.stride;
    }
    if (a.W < 2 * a.half + 8 || a.H < 2 * a.half + 8) return set_error(ctx, PMP_EINVAL, "image too small for the patch");
    if ((size_t)a.W * a.H * 4 > 0x7FFFFFFFull) return set_error(ctx, PMP_EINVAL, "image too large (4*W*H must be < 2^31)");
    a.agg_mode = p->agg_mode; a.edge_mode = p->edge_mode; a.coverage = p->coverage_mode;
    a.init_mode = p->init_mode; a.n_refine = p->n_refine;
    int k = p->topk < 1 ? 1 : p->topk;
    if (k > a.N) k = a.N;
    a.topk = k;
    a.inv_k = (float)(1.0 / (double)k);
    for (int v = 0; v < kMaxViews; ++v) a.rn[v] = (float)(1.0 / (double)(v + 1));
    Tables tb;
    rc = get_tables(ctx, a.W, a.H, p->sigma_color, tb);
    if (rc) return rc;
    ctx->last_sigma = p->sigma_color;
    a.ref = ir->second.gray;
    const int wrap = p->edge_mode == PMP_EDGE_WRAP ? 1 : 0;
    for (int v = 0; v < N; ++v) {
        auto iv = ctx->images.find(view_ids[v]);
        if (iv == ctx->images.end()) return set_error(ctx, PMP_ENOTFOUND, "unknown neighbour image id");
        if (iv->second.W != a.W || iv->second.H != a.H)
            return set_error(ctx, PMP_EINVAL, "neighbour images must have the reference's size");
        rc = ensure_quads(ctx, iv->second, wrap, stream);
        if (rc) return rc;
        a.vq[v] = iv->second.quads;
    }
    a.rowtab = tb.row; a.coltab = tb.col; a.Wc = tb.wc;
    derive_poses(ref_pose, a.N, view_poses, a.M, a.c);
    a.ku = (float)((double)a.W / (2.0 * M_PI));
    a.hu = (float)(0.5 * a.W);
    a.kv = (float)((double)a.H / M_PI);
    a.hv = (float)(0.5 * a.H);
    a.umax = (p->edge_mode == PMP_EDGE_WRAP) ? (float)a.W : (float)(a.W - 1);
    a.vmax = (float)(a.H - 1);
    a.var_floor = p->var_floor; a.fold_agree = p->fold_agree;
    a.depth_penalty = p->depth_penalty; a.plane_penalty = p->plane_penalty;
    float dr = p->refine_depth_range, nr = p->refine_normal_range;
    for (int t = 0; t < kMaxRefine; ++t) {
        a.dr[t] = dr; a.nr[t] = nr;
        dr = (float)((double)dr * (double)p->refine_depth_decay);
        nr = (float)((double)nr * (double)p->refine_normal_decay);
    }
    a.min_depth = p->min_depth; a.max_depth = p->max_depth;
    a.inv_lo = p->init_inv_depth_min; a.inv_hi = p->init_inv_depth_max;
    a.key0 = (uint32_t)(p->seed & 0xFFFFFFFFu);
    a.key1 = (uint32_t)(p->seed >> 32);
    return PMP_OK;
}

// The current keyframe as a one-slot descriptor set over the workspace, uploaded.
// dmask (device, nullable) is the finalize mask; with_mask: use the pooled
// mask of pmp_set_mask when dmask is null.  propagates: the call runs
// red-black passes from pass0, so the pass epochs are armed when
// pmp_debug_epochs has turned them on.
int plain_prepare(pmp_ctx* ctx, const pmp_params* p, const uint8_t* dmask, bool with_mask, int pass0,
                  bool propagates, QueueState*& q, QueueRun& r) {
    if (!ctx->views_set) return set_error(ctx, PMP_ESTATE, "pmp_set_views has not been called");
    int rc = queue_state(ctx, q);
    if (rc) return rc;
    if ((rc = desc_begin(ctx, q->plain))) return rc;
    pmp_keyframe kf;
    memset(&kf, 0, sizeof(kf));
    kf.ref_id = ctx->ref_id;
    memcpy(&kf.ref_pose, ctx->ref_pose, 12 * sizeof(double));
    kf.n_views = (int32_t)ctx->view_ids.size();
    for (int v = 0; v < kf.n_views; ++v) {
        kf.view_ids[v] = ctx->view_ids[v];
        memcpy(&kf.view_poses[v], &ctx->view_poses[12 * v], 12 * sizeof(double));
    }
    kf.mask_id = with_mask ? ctx->mask_id : -1;
    int items[3];
    rc = slot_fill(ctx, p, kf, ctx->ws, dmask, pass0, q->cfg.tile_rows, ctx->stream, q->plain.h_slots[0], items);
    if (rc) {
        if (rc == PMP_ENOTFOUND && with_mask && !dmask && ctx->mask_id >= 0 &&
            ctx->images.find((uint64_t)ctx->mask_id) == ctx->images.end())
            return set_error(ctx, PMP_ENOTFOUND, "mask image was dropped");
        return rc;
    }
    fill_list(q->plain.h_hdr->init, 1, &items[0]);
    fill_list(q->plain.h_hdr->prop, 1, &items[1]);
    fill_list(q->plain.h_hdr->rows, 1, &items[2]);
    if (propagates && (rc = epochs_arm(ctx, ctx->stream, 1, &q->plain.h_slots[0], &items[1], pass0))) return rc;
    if ((rc = queue_run(ctx, *q, q->cfg, p->edge_mode == PMP_EDGE_ZERO, p->postprocess != 0, r))) return rc;
    if (p->iterations > kMaxPasses / 2) return set_error(ctx, PMP_EINVAL, "iterations must be <= 64");
    return desc_upload(ctx, q->plain, 1, ctx->stream);
}

// A host mask goes to the workspace's staging buffer first.
// A host mask goes through a pinned copy, so the upload is a plain DMA that
// does not hold the host until the stream drains; a mask with the same bytes
// as the last one uploaded (a camera's fixed mask, or the same keyframe's mask
// in init, warm start and finalize) is not uploaded again: the device copy
// still holds it in stream order.
int stage_mask(pmp_ctx* ctx, const uint8_t* mask_host, const uint8_t*& dmask) {
    dmask = nullptr;
    if (!mask_host) return PMP_OK;
    const size_t P = (size_t)ctx->W * ctx->H;
    if (ctx->mask_held == P && memcmp(ctx->mask_pin, mask_host, P) == 0) {
        dmask = ctx->mask;
        return PMP_OK;
    }
    ctx->mask_held = 0;
    if (!ctx->mask_copied) PMP_HIP(ctx, hipEventCreateWithFlags(&ctx->mask_copied, hipEventDisableTiming));
    if (ctx->mask_pending) {
        PMP_HIP(ctx, hipEventSynchronize(ctx->mask_copied));   // the last upload out of mask_pin is done
        ctx->mask_pending = false;
    }
    if (P > ctx->mask_pin_cap) {
        if (ctx->mask_pin) (void)hipHostFree(ctx->mask_pin);
        ctx->mask_pin = nullptr;
        ctx->mask_pin_cap = 0;
        PMP_HIP(ctx, hipHostMalloc((void**)&ctx->mask_pin, P, hipHostMallocDefault));
        ctx->mask_pin_cap = P;
    }
    memcpy(ctx->mask_pin, mask_host, P);
    PMP_HIP(ctx, hipMemcpyAsync(ctx->mask, ctx->mask_pin, P, hipMemcpyHostToDevice, ctx->stream));
    PMP_HIP(ctx, hipEventRecord(ctx->mask_copied, ctx->stream));
    ctx->mask_pending = true;
    ctx->mask_held = P;
    dmask = ctx->mask;
    return PMP_OK;
}

}  // namespace pmp

namespace pmp {

// Unit camera-facing normal of a plane map: n = -p / |p| (p = n_c / d with
// p . b = 1/depth > 0 points away from the camera); (0, 0, 0) where p = 0.
__global__ void k_normals(const float4* __restrict__ plane, float* __restrict__ out, int P) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P) return;
    const float4 p = plane[k];
    const float s = fma_(p.z, p.z, fma_(p.y, p.y, p.x * p.x));
    float nx = 0.0f, ny = 0.0f, nz = 0.0f;
    if (s > 0.0f) {
        const float r = rsqrt_pos(med3_(s, 1e-30f, 1e30f));
        nx = -p.x * r;
        ny = -p.y * r;
        nz = -p.z * r;
    }
    out[3 * (size_t)k] = nx;
    out[3 * (size_t)k + 1] = ny;
    out[3 * (size_t)k + 2] = nz;
}

}  // namespace pmp

using namespace pmp;

extern "C" {

int pmp_set_views(pmp_ctx* ctx, uint64_t ref_id, const pmp_pose* ref_pose, int32_t n_views, const uint64_t* view_ids,
                  const pmp_pose* view_poses) {
    if (!ctx || !ref_pose || (n_views > 0 && (!view_ids || !view_poses))) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (n_views < 1 || n_views > kMaxViews) return set_error(ctx, PMP_EINVAL, "n_views must be in [1, 10]");
    auto it = ctx->images.find(ref_id);
    if (it == ctx->images.end()) return set_error(ctx, PMP_ENOTFOUND, "unknown reference image id");
    const int W = it->second.W, H = it->second.H;
    for (int v = 0; v < n_views; ++v) {
        auto iv = ctx->images.find(view_ids[v]);
        if (iv == ctx->images.end()) return set_error(ctx, PMP_ENOTFOUND, "unknown neighbour image id");
        if (iv->second.W != W || iv->second.H != H)
            return set_error(ctx, PMP_EINVAL, "neighbour images must have the reference's size");
    }
    int rc = ensure_workspace(ctx, W, H);
    if (rc) return rc;
    ctx->ref_id = ref_id;
    ctx->view_ids.assign(view_ids, view_ids + n_views);
    memcpy(ctx->ref_pose, ref_pose->R, 9 * sizeof(double));
    memcpy(ctx->ref_pose + 9, ref_pose->t, 3 * sizeof(double));
    ctx->view_poses.resize(12 * (size_t)n_views);
    for (int v = 0; v < n_views; ++v) {
        memcpy(&ctx->view_poses[12 * v], view_poses[v].R, 9 * sizeof(double));
        memcpy(&ctx->view_poses[12 * v + 9], view_poses[v].t, 3 * sizeof(double));
    }
    ctx->W = W; ctx->H = H;
    ctx->views_set = true;
    return PMP_OK;
}

int pmp_set_mask(pmp_ctx* ctx, int64_t mask_id) {
    if (!ctx) return PMP_EINVAL;
    if (mask_id >= 0 && ctx->images.find((uint64_t)mask_id) == ctx->images.end())
        return set_error(ctx, PMP_ENOTFOUND, "unknown mask image id");
    ctx->mask_id = mask_id < 0 ? -1 : mask_id;
    return PMP_OK;
}

int pmp_random_initialization(pmp_ctx* ctx, const pmp_params* params) {
    TraceRange trace_range("pmp_random_initialization");
    if (!ctx || !params) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    QueueState* q = nullptr;
    QueueRun r;
    int rc = plain_prepare(ctx, params, nullptr, false, 0, false, q, r);
    if (rc) return rc;
    const int tp = timing_begin(ctx, kStageInit);
    rc = q_init(ctx, ctx->stream, q->plain, r);
    timing_end(ctx, tp);
    return rc;
}

int pmp_score_plane_depth(pmp_ctx* ctx, const pmp_params* params) {
    TraceRange trace_range("pmp_score_plane_depth");
    if (!ctx || !params) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    QueueState* q = nullptr;
    QueueRun r;
    int rc = plain_prepare(ctx, params, nullptr, false, 0, false, q, r);
    if (rc) return rc;
    return q_score(ctx, ctx->stream, q->plain, r);
}

int pmp_red_black_pass(pmp_ctx* ctx, const pmp_params* params, int32_t it_begin, int32_t it_end) {
    TraceRange trace_range("pmp_red_black_pass");
    if (!ctx || !params || it_begin < 0 || it_end < it_begin) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (it_end - it_begin > kMaxPasses / 2) return set_error(ctx, PMP_EINVAL, "at most 64 iterations per call");
    QueueState* q = nullptr;
    QueueRun r;
    int rc = plain_prepare(ctx, params, nullptr, false, 2 * it_begin, true, q, r);
    if (rc) return rc;
    const int passes = 2 * (it_end - it_begin);
    if ((rc = q_clear(ctx, ctx->stream, q->plain, passes))) return rc;
    for (int pass = 0; pass < passes; ++pass) {
        const int tp = timing_begin(ctx, kStagePropagate);
        rc = q_prop(ctx, ctx->stream, q->plain, r, pass);
        timing_end(ctx, tp);
        if (rc) return rc;
    }
    return PMP_OK;
}

int pmp_finalize(pmp_ctx* ctx, const pmp_params* params, const uint8_t* mask) {
    TraceRange trace_range("pmp_finalize");
    if (!ctx || !params) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (!ctx->views_set) return set_error(ctx, PMP_ESTATE, "pmp_set_views has not been called");
    const uint8_t* dmask = nullptr;
    int rc = stage_mask(ctx, mask, dmask);
    if (rc) return rc;
    QueueState* q = nullptr;
    QueueRun r;
    if ((rc = plain_prepare(ctx, params, dmask, true, 0, false, q, r))) return rc;
    const int tp = timing_begin(ctx, kStageFinalize);
    rc = q_finalize(ctx, ctx->stream, q->plain, r);
    timing_end(ctx, tp);
    return rc;
}

int pmp_densify(pmp_ctx* ctx, const pmp_params* params, const uint8_t* mask) {
    TraceRange trace_range("pmp_densify");
    if (!ctx || !params || params->iterations < 0) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (!ctx->views_set) return set_error(ctx, PMP_ESTATE, "pmp_set_views has not been called");
    const uint8_t* dmask = nullptr;
    int rc = stage_mask(ctx, mask, dmask);
    if (rc) return rc;
    QueueState* q = nullptr;
    QueueRun r;
    if ((rc = plain_prepare(ctx, params, dmask, true, 0, true, q, r))) return rc;
    if ((rc = q_densify(ctx, ctx->stream, q->plain, r, 2 * params->iterations, q->cfg.use_graph != 0))) return rc;
    ctx->densify_calls++;
    return PMP_OK;
}

int pmp_set_maps(pmp_ctx* ctx, const float* plane4, const float* depth, const float* score) {
    if (!ctx) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (!ctx->views_set) return set_error(ctx, PMP_ESTATE, "pmp_set_views has not been called");
    const size_t P = (size_t)ctx->W * ctx->H;
    if (plane4) PMP_HIP(ctx, hipMemcpyAsync(ctx->ws.plane, plane4, P * 16, hipMemcpyHostToDevice, ctx->stream));
    if (depth) PMP_HIP(ctx, hipMemcpyAsync(ctx->ws.depth, depth, P * 4, hipMemcpyHostToDevice, ctx->stream));
    if (score) PMP_HIP(ctx, hipMemcpyAsync(ctx->ws.score, score, P * 4, hipMemcpyHostToDevice, ctx->stream));
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PMP_OK;
}

int pmp_get_maps(pmp_ctx* ctx, int32_t which, float* plane4, float* depth, float* score, int32_t* nghbr) {
    if (!ctx) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (!ctx->views_set) return set_error(ctx, PMP_ESTATE, "pmp_set_views has not been called");
    const size_t P = (size_t)ctx->W * ctx->H;
    const void *sp, *sd, *ss;
    if (which == PMP_MAPS_WORK) { sp = ctx->ws.plane; sd = ctx->ws.depth; ss = ctx->ws.score; }
    else if (which == PMP_MAPS_INIT) { sp = ctx->ws.iplane; sd = ctx->ws.idepth; ss = ctx->ws.iscore; }
    else if (which == PMP_MAPS_FINAL) { sp = ctx->ws.oplane; sd = ctx->ws.odepth; ss = ctx->ws.oscore; }
    else return set_error(ctx, PMP_EINVAL, "unknown map set");
    if (plane4) PMP_HIP(ctx, hipMemcpyAsync(plane4, sp, P * 16, hipMemcpyDeviceToHost, ctx->stream));
    if (depth) PMP_HIP(ctx, hipMemcpyAsync(depth, sd, P * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (score) PMP_HIP(ctx, hipMemcpyAsync(score, ss, P * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (nghbr) {
        if (which != PMP_MAPS_FINAL) return set_error(ctx, PMP_EINVAL, "nghbr exists only for the final maps");
        PMP_HIP(ctx, hipMemcpyAsync(nghbr, ctx->ws.onghbr, P * 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PMP_OK;
}

int pmp_get_normals(pmp_ctx* ctx, int32_t which, float* normal3) {
    if (!ctx || !normal3) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (!ctx->views_set) return set_error(ctx, PMP_ESTATE, "pmp_set_views has not been called");
    const size_t P = (size_t)ctx->W * ctx->H;
    const float4* sp;
    if (which == PMP_MAPS_WORK) sp = (const float4*)ctx->ws.plane;
    else if (which == PMP_MAPS_INIT) sp = (const float4*)ctx->ws.iplane;
    else if (which == PMP_MAPS_FINAL) sp = (const float4*)ctx->ws.oplane;
    else return set_error(ctx, PMP_EINVAL, "unknown map set");
    if (ctx->normals_cap < 3 * P) {
        if (ctx->normals) (void)hipFree(ctx->normals);
        ctx->normals = nullptr;
        ctx->normals_cap = 0;
        PMP_HIP(ctx, hipMalloc(&ctx->normals, 3 * P * sizeof(float)));
        ctx->normals_cap = 3 * P;
    }
    hipLaunchKernelGGL(k_normals, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, ctx->stream, sp, ctx->normals, (int)P);
    PMP_HIP(ctx, hipGetLastError());
    PMP_HIP(ctx, hipMemcpyAsync(normal3, ctx->normals, 3 * P * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PMP_OK;
}

int pmp_eval_counts(pmp_ctx* ctx, int64_t* evals, int64_t* candidates) {
    if (!ctx) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    unsigned long long h[2] = {0, 0};
    if (ctx->counters) {
        PMP_HIP(ctx, hipMemcpyAsync(h, ctx->counters, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (evals) *evals = (int64_t)h[0];
    if (candidates) *candidates = (int64_t)h[1];
    return PMP_OK;
}

}  // extern "C"
