// pmp_context.hip -- context, errors, image pool, workspace and host tables.
#include <math.h>
#include <string.h>

#include <string>
#include <vector>

#include "pmp_trace.h"
#include "pmp_internal.h"

namespace pmp {

int set_error(pmp_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

int upload_host_rows(pmp_ctx* ctx, void* dst, const uint8_t* src, size_t row, size_t rows, size_t pitch,
                     hipStream_t s) {
    const size_t bytes = row * rows;
    if (bytes == 0) return PMP_OK;
    if (bytes > ctx->stage_cap) {
        if (ctx->stage) (void)hipHostFree(ctx->stage);
        ctx->stage = nullptr;
        ctx->stage_cap = 0;
        PMP_HIP(ctx, hipHostMalloc((void**)&ctx->stage, bytes, hipHostMallocDefault));
        ctx->stage_cap = bytes;
    }
    if (pitch == row) {
        memcpy(ctx->stage, src, bytes);
    } else {
        for (size_t r = 0; r < rows; ++r) memcpy(ctx->stage + r * row, src + r * pitch, row);
    }
    PMP_HIP(ctx, hipMemcpyAsync(dst, ctx->stage, bytes, hipMemcpyHostToDevice, s));
    // the staging buffer is reused by the next upload
    PMP_HIP(ctx, hipStreamSynchronize(s));
    return PMP_OK;
}

int check_hip(pmp_ctx* ctx, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_error(ctx, e == hipErrorOutOfMemory ? PMP_ENOMEM : PMP_EDEVICE, m);
}

// Bearing tables (PatchMatchCuda.cu:191-198), double precision rounded once:
// rowtab[i] = (cos lat_i, sin lat_i), lat_i = -pi (i/H - 1/2);
// coltab[j] = (sin lon_j, cos lon_j), lon_j = 2 pi (j/W - 1/2).
float2 table_row(int i, int H) {
    const double lat = -M_PI * ((double)i / (double)H - 0.5);
    return make_float2((float)cos(lat), (float)sin(lat));
}

float2 table_col(int j, int W) {
    const double lon = 2.0 * M_PI * ((double)j / (double)W - 0.5);
    return make_float2((float)sin(lon), (float)cos(lon));
}

void host_tables(int W, int H, std::vector<float2>& rowtab, std::vector<float2>& coltab) {
    rowtab.resize(H);
    coltab.resize(W);
    for (int i = 0; i < H; ++i) rowtab[i] = table_row(i, H);
    for (int j = 0; j < W; ++j) coltab[j] = table_col(j, W);
}

// M = R_n R_ref^T, c = t_n - M t_ref in double, rounded once.  Equivalent to
// the two transforms of PatchMatchCuda.cu:124-130 with R1_inv = R_ref^T,
// t1_inv = -R_ref^T t_ref (depthmap.cc:604-628) and the neighbours' own
// world->camera poses (depthmap.cc:631-649).
void derive_poses(const double* ref_pose, int N, const double* view_poses, float (*M)[9], float (*c)[3]) {
    const double* R0 = ref_pose;
    const double* t0 = ref_pose + 9;
    for (int n = 0; n < N; ++n) {
        const double* Rn = view_poses + 12 * n;
        const double* tn = Rn + 9;
        double Md[9];
        for (int r = 0; r < 3; ++r)
            for (int col = 0; col < 3; ++col) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k) s += Rn[3 * r + k] * R0[3 * col + k];
                Md[3 * r + col] = s;
            }
        for (int r = 0; r < 3; ++r) {
            double s = tn[r];
            for (int k = 0; k < 3; ++k) s -= Md[3 * r + k] * t0[k];
            c[n][r] = (float)s;
        }
        for (int k = 0; k < 9; ++k) M[n][k] = (float)Md[k];
    }
}

void free_slot(SlotBuf& b) {
    void* p[] = {b.plane, b.iplane, b.oplane, b.depth, b.score, b.idepth, b.iscore, b.odepth, b.oscore,
                 b.tmp0, b.tmp1, b.onghbr};
    for (void* x : p)
        if (x) (void)hipFree(x);
    b = SlotBuf();
}

template <typename T>
static int alloc_n(pmp_ctx* ctx, T*& p, size_t n) {
    PMP_HIP(ctx, hipMalloc(&p, n * sizeof(T)));
    return PMP_OK;
}

// Sizes b for a W x H keyframe; reallocates (after the caller has made sure no
// queued work uses the old buffers) only when it grows.
int ensure_slot(pmp_ctx* ctx, SlotBuf& b, int W, int H) {
    const size_t P = (size_t)W * H;
    if (P > b.cap) {
        free_slot(b);
        int rc = 0;
        rc |= alloc_n(ctx, b.plane, P);
        rc |= alloc_n(ctx, b.iplane, P);
        rc |= alloc_n(ctx, b.oplane, P);
        rc |= alloc_n(ctx, b.depth, P);
        rc |= alloc_n(ctx, b.score, P);
        rc |= alloc_n(ctx, b.idepth, P);
        rc |= alloc_n(ctx, b.iscore, P);
        rc |= alloc_n(ctx, b.odepth, P);
        rc |= alloc_n(ctx, b.oscore, P);
        rc |= alloc_n(ctx, b.tmp0, P);
        rc |= alloc_n(ctx, b.tmp1, P);
        rc |= alloc_n(ctx, b.onghbr, P);
        if (rc) {
            free_slot(b);
            return rc;
        }
        b.cap = P;
    }
    b.W = W;
    b.H = H;
    return PMP_OK;
}

int ensure_workspace(pmp_ctx* ctx, int W, int H) {
    const size_t P = (size_t)W * H;
    if (P > ctx->ws.cap) {
        PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->mask) (void)hipFree(ctx->mask);
        ctx->mask = nullptr;
        ctx->mask_held = 0;
        int rc = ensure_slot(ctx, ctx->ws, W, H);
        if (rc) return rc;
        if ((rc = alloc_n(ctx, ctx->mask, P))) return rc;
    }
    ctx->ws.W = W;
    ctx->ws.H = H;
    if (!ctx->counters) {
        PMP_HIP(ctx, hipMalloc(&ctx->counters, 4 * sizeof(unsigned long long)));
        PMP_HIP(ctx, hipMemsetAsync(ctx->counters, 0, 4 * sizeof(unsigned long long), ctx->stream));
    }
    return PMP_OK;
}

int get_tables(pmp_ctx* ctx, int W, int H, float sigma_color, Tables& t) {
    auto ir = ctx->rowtabs.find(H);
    auto ic = ctx->coltabs.find(W);
    if (ir == ctx->rowtabs.end() || ic == ctx->coltabs.end()) {
        std::vector<float2> rt, ct;
        host_tables(W, H, rt, ct);
        if (ir == ctx->rowtabs.end()) {
            float2* d = nullptr;
            PMP_HIP(ctx, hipMalloc(&d, H * sizeof(float2)));
            ctx->rowtabs[H] = d;
            PMP_HIP(ctx, hipMemcpy(d, rt.data(), H * sizeof(float2), hipMemcpyHostToDevice));
            ir = ctx->rowtabs.find(H);
        }
        if (ic == ctx->coltabs.end()) {
            float2* d = nullptr;
            PMP_HIP(ctx, hipMalloc(&d, W * sizeof(float2)));
            ctx->coltabs[W] = d;
            PMP_HIP(ctx, hipMemcpy(d, ct.data(), W * sizeof(float2), hipMemcpyHostToDevice));
            ic = ctx->coltabs.find(W);
        }
    }
    uint32_t key;
    memcpy(&key, &sigma_color, 4);
    auto iw = ctx->wcs.find(key);
    if (iw == ctx->wcs.end()) {
        float h[256];
        const double cf = 1.0 / (2.0 * (double)sigma_color * (double)sigma_color);
        for (int a = 0; a < 256; ++a) h[a] = (float)exp(-(double)(a * a) * cf);
        float* d = nullptr;
        PMP_HIP(ctx, hipMalloc(&d, sizeof(h)));
        ctx->wcs[key] = d;
        PMP_HIP(ctx, hipMemcpy(d, h, sizeof(h), hipMemcpyHostToDevice));
        iw = ctx->wcs.find(key);
    }
    t.row = ir->second;
    t.col = ic->second;
    t.wc = iw->second;
    return PMP_OK;
}

// Column pairs: halfword (x, y) holds p(x, y) | p(x, y+1) << 8 in rows of
// W + 2 halfwords.  Columns W and W+1 repeat columns 0 and 1 (longitude wrap:
// a sample at u = W reads them with no index wrap) or column W-1 (clamp); y+1
// of the last row repeats it.  A bilinear sample at (x, y), x in [0, W], is
// the 4 bytes at halfword (x, y): p(x,y), p(x,y+1), p(x+1,y), p(x+1,y+1).
__global__ void k_pack(const uint8_t* __restrict__ g, uint32_t* __restrict__ q32, int W, int H, int wrap) {
    uint16_t* q = (uint16_t*)q32;
    int xo = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (xo > W + 1 || y >= H) return;
    int x = xo < W ? xo : (wrap ? xo - W : W - 1);
    int y1 = y + 1 < H ? y + 1 : H - 1;
    q[(size_t)y * (W + 2) + xo] = (uint16_t)(g[(size_t)y * W + x] | (g[(size_t)y1 * W + x] << 8));
}

int ensure_quads(pmp_ctx* ctx, Image& im, int wrap, hipStream_t stream) {
    if (im.quads && im.quads_wrap == wrap) return PMP_OK;
    if (!im.quads) PMP_HIP(ctx, hipMalloc(&im.quads, Image::quad_bytes(im.W, im.H)));
    dim3 blk(256), grd((im.W + 2 + 255) / 256, im.H);
    hipLaunchKernelGGL(k_pack, grd, blk, 0, stream, im.gray, im.quads, im.W, im.H, wrap);
    PMP_HIP(ctx, hipGetLastError());
    im.quads_wrap = wrap;
    return PMP_OK;
}

int timing_begin(pmp_ctx* ctx, int stage) {
    if (!ctx->timing) return -1;
    const size_t k = ctx->events_used;
    while (ctx->events.size() < 2 * (k + 1)) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        ctx->events.push_back(e);
    }
    if (ctx->event_stage.size() < k + 1) ctx->event_stage.resize(k + 1);
    ctx->event_stage[k] = stage;
    if (hipEventRecord(ctx->events[2 * k], ctx->stream) != hipSuccess) return -1;
    ctx->events_used = k + 1;
    return (int)k;
}

void timing_end(pmp_ctx* ctx, int pair) {
    if (pair < 0) return;
    (void)hipEventRecord(ctx->events[2 * (size_t)pair + 1], ctx->stream);
}

static void free_image(Image& im) {
    if (im.gray) (void)hipFree(im.gray);
    if (im.quads) (void)hipFree(im.quads);
    im = Image();
}

}  // namespace pmp

using namespace pmp;

extern "C" {

int pmp_create(pmp_ctx** out, int32_t device) {
    if (!out) return PMP_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return PMP_EDEVICE;
    if (device < 0 || device >= n) return PMP_EINVAL;
    pmp_ctx* ctx = new (std::nothrow) pmp_ctx();
    if (!ctx) return PMP_ENOMEM;
    ctx->device = device;
    DeviceScope dev(device);
    if (!dev.ok() || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return PMP_EDEVICE;
    }
    *out = ctx;
    return PMP_OK;
}

void pmp_destroy(pmp_ctx* ctx) {
    if (!ctx) return;
    DeviceScope dev(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    queue_free(ctx);
    preinit_free(ctx);
    free_slot(ctx->ws);
    for (auto& kv : ctx->images) free_image(kv.second);
    void* bufs[] = {ctx->mask, ctx->counters, ctx->thumb.src, ctx->thumb.gray, ctx->thumb.dst, ctx->thumb.taps,
                    ctx->clean_maps, ctx->clean_poses, ctx->normals, ctx->prune_z, ctx->prune_alive, ctx->prune_out,
                    ctx->prune_pts, ctx->prune_cols, ctx->prune_kps, ctx->prune_blk, ctx->prune_poses, ctx->dbg,
                    ctx->epochs};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (ctx->stage) (void)hipHostFree(ctx->stage);
    if (ctx->mask_pin) (void)hipHostFree(ctx->mask_pin);
    if (ctx->mask_copied) (void)hipEventDestroy(ctx->mask_copied);
    for (auto& kv : ctx->rowtabs) (void)hipFree(kv.second);
    for (auto& kv : ctx->coltabs) (void)hipFree(kv.second);
    for (auto& kv : ctx->wcs) (void)hipFree(kv.second);
    for (auto& v : ctx->pruner) {
        if (v.depth) (void)hipFree(v.depth);
        if (v.bgr) (void)hipFree(v.bgr);
    }
    for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* pmp_last_error(const pmp_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

void pmp_default_params(pmp_params* p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->patch_half = 5;
    p->patch_stride = 2;
    p->iterations = 2;
    p->n_refine = 6;
    p->agg_mode = PMP_AGG_REF_FOLD;
    p->topk = 2;
    p->edge_mode = PMP_EDGE_WRAP;
    p->coverage_mode = PMP_COVER_REF_GRID;
    p->init_mode = PMP_INIT_REF_UNIFORM;
    p->postprocess = 1;
    p->seed = 0;
    p->min_depth = 1e-5f;
    p->max_depth = 10.0f;
    p->init_inv_depth_min = 1.0f / 20.0f;
    p->init_inv_depth_max = 1.0f / 0.3f;
    p->min_score = 0.1f;
    p->sigma_color = 50.0f;
    p->sigma_space = 5.0f;
    p->var_floor = 0.1f;
    p->fold_agree = 0.25f;
    p->depth_penalty = 1250.0f;
    p->plane_penalty = 8.0f;
    p->refine_depth_range = 0.02f;
    p->refine_normal_range = 0.5f;
    p->refine_depth_decay = 0.3f;
    p->refine_normal_decay = 0.8f;
    p->min_patch_sd = 0.0f;
    p->median_rel_thr = 0.05;
}

int pmp_synchronize(pmp_ctx* ctx) {
    if (!ctx) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PMP_OK;
}

int pmp_set_timing(pmp_ctx* ctx, int32_t enable) {
    if (!ctx) return PMP_EINVAL;
    ctx->timing = enable != 0;
    ctx->events_used = 0;
    ctx->densify_calls = 0;
    return PMP_OK;
}

int pmp_get_timing(pmp_ctx* ctx, pmp_timing* out) {
    if (!ctx || !out) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    memset(out, 0, sizeof(*out));
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t k = 0; k < ctx->events_used; ++k) {
        float ms = 0.0f;
        PMP_HIP(ctx, hipEventElapsedTime(&ms, ctx->events[2 * k], ctx->events[2 * k + 1]));
        switch (ctx->event_stage[k]) {
            case kStageInit: out->init_ms += ms; break;
            case kStagePropagate: out->propagate_ms += ms; out->propagate_launches++; break;
            default: out->finalize_ms += ms; break;
        }
    }
    out->densify_calls = ctx->densify_calls;
    ctx->events_used = 0;
    ctx->densify_calls = 0;
    return PMP_OK;
}

int pmp_image_put(pmp_ctx* ctx, uint64_t id, const uint8_t* pixels, int32_t width, int32_t height, int64_t pitch,
                  int32_t on_device) {
    TraceRange trace_range("pmp_image_put");
    if (!ctx || !pixels || width <= 0 || height <= 0 || pitch < width) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    auto it = ctx->images.find(id);
    if (it != ctx->images.end()) {
        PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        free_image(it->second);
        ctx->images.erase(it);
    }
    Image im;
    im.W = width;
    im.H = height;
    const size_t bytes = (size_t)width * height;
    PMP_HIP(ctx, hipMalloc(&im.gray, bytes));
    im.bytes = bytes;
    int rc;
    if (on_device) {
        hipError_t e = hipMemcpy2DAsync(im.gray, width, pixels, (size_t)pitch, width, height, hipMemcpyDeviceToDevice,
                                        ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        rc = e == hipSuccess ? PMP_OK : check_hip(ctx, e, "pmp_image_put copy");
    } else {
        rc = upload_host_rows(ctx, im.gray, pixels, (size_t)width, (size_t)height, (size_t)pitch, ctx->stream);
    }
    if (rc) {
        free_image(im);
        return rc;
    }
    ctx->images[id] = im;
    return PMP_OK;
}

int pmp_image_get(pmp_ctx* ctx, uint64_t id, uint8_t* pixels) {
    if (!ctx || !pixels) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    auto it = ctx->images.find(id);
    if (it == ctx->images.end()) return set_error(ctx, PMP_ENOTFOUND, "unknown image id");
    PMP_HIP(ctx, hipMemcpyAsync(pixels, it->second.gray, it->second.bytes, hipMemcpyDeviceToHost, ctx->stream));
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PMP_OK;
}

int pmp_image_drop(pmp_ctx* ctx, uint64_t id) {
    if (!ctx) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    auto it = ctx->images.find(id);
    if (it == ctx->images.end()) return set_error(ctx, PMP_ENOTFOUND, "unknown image id");
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    free_image(it->second);
    ctx->images.erase(it);
    if (ctx->views_set) {
        bool used = ctx->ref_id == id;
        for (uint64_t v : ctx->view_ids) used = used || v == id;
        if (used) ctx->views_set = false;
    }
    return PMP_OK;
}

int pmp_image_size(const pmp_ctx* ctx, uint64_t id, int32_t* width, int32_t* height) {
    if (!ctx) return PMP_EINVAL;
    auto it = ctx->images.find(id);
    if (it == ctx->images.end()) return PMP_ENOTFOUND;
    if (width) *width = it->second.W;
    if (height) *height = it->second.H;
    return PMP_OK;
}

int pmp_image_pool_stats(const pmp_ctx* ctx, int64_t* count, int64_t* bytes) {
    if (!ctx) return PMP_EINVAL;
    int64_t b = 0;
    for (auto& kv : ctx->images) b += (int64_t)kv.second.pool_bytes();
    if (count) *count = (int64_t)ctx->images.size();
    if (bytes) *bytes = b;
    return PMP_OK;
}

}  // extern "C"
