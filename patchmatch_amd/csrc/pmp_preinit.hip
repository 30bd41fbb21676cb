// pmp_preinit.hip -- the host pre-initialisation of ComputePatchMatch
// (depthmap.cc:690-691) on the device, between the random initialisation and
// the red-black passes:
//   PreInitSparsePCL2 (:361-432)   every sparse landmark seeds the pixel it
//                                  reprojects to with a camera-facing plane at
//                                  its distance;
//   PreInitResultWarp (:434-516)   every pixel of the memorised previous
//                                  result (MemorizeResult, :808-829) is moved
//                                  along its plane into the current camera.
// A candidate replaces a pixel's hypothesis when its score beats the pixel's.
// The reference runs both loops with OpenMP over shared maps (a race when two
// candidates land on one pixel); here the outcome is that of the sequential
// loop: per pixel the first candidate, in loop order, with the highest score
// wins if it beats the pixel's score.  Three launches per stage:
//   k_pre_*       candidate target, plane, depth and score; a 64-bit
//                 atomicMax of (ordered score, ~index) per target pixel picks
//                 the winner;
//   k_pre_apply   the winner writes the working maps and the copy-back
//                 snapshot (the reference's result maps).
#include <math.h>
#include <string.h>

#include "pmp_geo.h"
#include "pmp_trace.h"
#include "pmp_internal.h"
#include "pmp_math.h"
#include "pmp_queue.h"
#include "pmp_score.h"
#include "pmp_tile.h"

#define PMP_CONST __attribute__((address_space(4)))

namespace pmp {

struct PreArgs {
    double pose[12];          // current reference, world->camera
    double prev_pose[12];     // previous reference (warp)
    int n;                    // candidates
};

struct PreState {
    size_t cap = 0;           // candidates
    int* cpix = nullptr;
    float4* cplane = nullptr;
    float* cdep = nullptr;
    float* cscore = nullptr;
    size_t keys_cap = 0;      // pixels
    unsigned long long* keys = nullptr;
    size_t xyz_cap = 0;
    double* xyz = nullptr;
    unsigned long long* accepted = nullptr;
    size_t prev_cap = 0;
    float4* prev_plane = nullptr;
    float* prev_depth = nullptr;
    int prev_W = 0, prev_H = 0;
    double prev_pose[12];
    bool prev_valid = false;
};

__device__ __forceinline__ const SlotDesc& slot0(const SlotDesc PMP_CONST* slots) {
    return *(const SlotDesc*)slots;
}

__device__ __forceinline__ GeoArgs geo_of(const KfArgs& a) {
    GeoArgs g;
    g.W = a.W; g.H = a.H; g.rowtab = a.rowtab; g.coltab = a.coltab;
    g.ku = a.ku; g.hu = a.hu; g.kv = a.kv; g.hv = a.hv;
    return g;
}

// Orders candidates by score, then by lower index (the sequential loop keeps
// the first of equal scores); -0 is folded into +0, which compares equal.
__device__ __forceinline__ unsigned long long cand_key(float score, int q) {
    const float s = score + 0.0f;
    const unsigned int u = __float_as_uint(s);
    const unsigned int o = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return ((unsigned long long)o << 32) | (unsigned long long)(0xFFFFFFFFu - (unsigned int)q);
}

// PlanePosFromDepthAndNormal_equirect (depthmap.cc:130-141): n / (n.(d b)),
// the denominator pushed away from 0 to +-1e-5
__device__ __forceinline__ void plane_eps(float bx, float by, float bz, float dep, float nx, float ny, float nz,
                                          float& p0, float& p1, float& p2) {
    const float qx = dep * bx, qy = dep * by, qz = dep * bz;
    float den = fma_(nz, qz, fma_(ny, qy, nx * qx));
    if (den < 1e-5f && den >= 0.0f) den = 1e-5f;
    else if (den > -1e-5f && den < 0.0f) den = -1e-5f;
    const float r = rcp(den);
    p0 = nx * r;
    p1 = ny * r;
    p2 = nz * r;
}

// Target pixel of a world point (ReprojectEquirect, then int(x + 0.5),
// int(y + 0.5)): -1 when too near, outside the image or masked.
__device__ __forceinline__ int target_pixel(const GeoArgs& g, const double* pose, double X0, double X1, double X2,
                                            const uint8_t* mask, float& dist) {
    float x, y;
    if (reproject(g, pose, X0, X1, X2, x, y, dist)) return -1;
    const int j = (int)(x + 0.5f), i = (int)(y + 0.5f);
    if (i < 0 || j < 0 || i >= g.H || j >= g.W) return -1;
    const int k = i * g.W + j;
    if (mask && !mask[k]) return -1;
    return k;
}

// PosPatchPartOnPlane (depthmap.cc:163-177): scale = p.b in float pushed away
// from 0 to +-1e-6, b/scale in float, R^T (X - t) in double.
__device__ __forceinline__ void pos_on_plane(const GeoArgs& g, const double* pose, int i, int j, float p0, float p1,
                                             float p2, double& X0, double& X1, double& X2) {
    const float2 r = g.rowtab[i], c = g.coltab[j];
    const float bx = r.x * c.x, by = -r.y, bz = r.x * c.y;
    float s = fma_(p2, bz, fma_(p1, by, p0 * bx));
    if (s < 1e-6f && s >= 0.0f) s = 1e-6f;
    else if (s > -1e-6f && s < 0.0f) s = -1e-6f;
    const float rs = rcp(s);
    const double q0 = (double)(bx * rs) - pose[9], q1 = (double)(by * rs) - pose[10], q2 = (double)(bz * rs) - pose[11];
    X0 = pose[0] * q0 + pose[3] * q1 + pose[6] * q2;
    X1 = pose[1] * q0 + pose[4] * q1 + pose[7] * q2;
    X2 = pose[2] * q0 + pose[5] * q1 + pose[8] * q2;
}

// Scores the candidate (pix, plane) and offers it to its pixel.
template <bool kZero>
__device__ __forceinline__ void offer(const SlotDesc& d, const float* sWc, float* sv, int q, int pix, float p0,
                                      float p1, float p2, float dist, float4* __restrict__ cplane,
                                      float* __restrict__ cdep, float* __restrict__ cscore,
                                      unsigned long long* __restrict__ keys) {
    const int i = pix / d.a.W, j = pix - (pix / d.a.W) * d.a.W;
    const RefStat rs = ref_stats(d.a, sWc, i, j);
    const float sc = score_plane<0, kZero>(d.a, sWc, rs, i, j, p0, p1, p2, sv, 256);
    cplane[q] = make_float4(p0, p1, p2, 0.0f);
    cdep[q] = dist;
    cscore[q] = sc;
    atomicMax(&keys[pix], cand_key(sc, q));
}

template <bool kZero>
__global__ __launch_bounds__(256) void k_pre_landmarks(const SlotDesc PMP_CONST* slots, PreArgs pa,
                                                       const double* __restrict__ xyz, int* __restrict__ cpix,
                                                       float4* __restrict__ cplane, float* __restrict__ cdep,
                                                       float* __restrict__ cscore,
                                                       unsigned long long* __restrict__ keys) {
    __shared__ float sWc[256];
    __shared__ float sv[kMaxViews][256];
    const SlotDesc& d = slot0(slots);
    load_wc(d.a.Wc, sWc);
    const GeoArgs g = geo_of(d.a);
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < pa.n; q += gridDim.x * blockDim.x) {
        float dist;
        int pix = target_pixel(g, pa.pose, xyz[3 * q], xyz[3 * q + 1], xyz[3 * q + 2], d.mask, dist);
        if (pix >= 0 && !in_score_grid(d.a, pix / g.W, pix % g.W)) pix = -1;
        cpix[q] = pix;
        if (pix < 0) continue;
        float bx, by, bz;
        bearing(d.a, pix / g.W, pix % g.W, bx, by, bz);
        float p0, p1, p2;
        plane_eps(bx, by, bz, dist, -bx, -by, -bz, p0, p1, p2);   // normal: -bearing (:393-395)
        offer<kZero>(d, sWc, &sv[0][threadIdx.x], q, pix, p0, p1, p2, dist, cplane, cdep, cscore, keys);
    }
}

template <bool kZero>
__global__ __launch_bounds__(256) void k_pre_warp(const SlotDesc PMP_CONST* slots, PreArgs pa,
                                                  const float4* __restrict__ prev_plane,
                                                  const float* __restrict__ prev_depth, int* __restrict__ cpix,
                                                  float4* __restrict__ cplane, float* __restrict__ cdep,
                                                  float* __restrict__ cscore, unsigned long long* __restrict__ keys) {
    __shared__ float sWc[256];
    __shared__ float sv[kMaxViews][256];
    const SlotDesc& d = slot0(slots);
    load_wc(d.a.Wc, sWc);
    const GeoArgs g = geo_of(d.a);
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < pa.n; q += gridDim.x * blockDim.x) {
        cpix[q] = -1;
        // the current mask at the previous pixel (:445), depth >= 1e-4 (:450)
        if (d.mask && !d.mask[q]) continue;
        if (!(prev_depth[q] >= 1e-4f)) continue;
        const float4 pp = prev_plane[q];
        double X0, X1, X2;
        pos_on_plane(g, pa.prev_pose, q / g.W, q % g.W, pp.x, pp.y, pp.z, X0, X1, X2);
        float dist;
        const int pix = target_pixel(g, pa.pose, X0, X1, X2, d.mask, dist);
        if (pix < 0 || !in_score_grid(d.a, pix / g.W, pix % g.W)) continue;
        cpix[q] = pix;
        // normal: the previous plane's direction, not rotated (:479-481)
        const float nr2 = clampf(fma_(pp.z, pp.z, fma_(pp.y, pp.y, pp.x * pp.x)), 1e-30f, 1e30f);
        const float inr = rsqrt_pos(nr2);
        float bx, by, bz;
        bearing(d.a, pix / g.W, pix % g.W, bx, by, bz);
        float p0, p1, p2;
        plane_eps(bx, by, bz, dist, pp.x * inr, pp.y * inr, pp.z * inr, p0, p1, p2);
        offer<kZero>(d, sWc, &sv[0][threadIdx.x], q, pix, p0, p1, p2, dist, cplane, cdep, cscore, keys);
    }
}

__global__ __launch_bounds__(256) void k_pre_apply(const SlotDesc PMP_CONST* slots, int n,
                                                   const int* __restrict__ cpix, const float4* __restrict__ cplane,
                                                   const float* __restrict__ cdep, const float* __restrict__ cscore,
                                                   const unsigned long long* __restrict__ keys,
                                                   unsigned long long* __restrict__ accepted) {
    const SlotDesc& d = slot0(slots);
    unsigned int acc = 0;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int pix = cpix[q];
        if (pix < 0) continue;
        const float sc = cscore[q];
        if (keys[pix] != cand_key(sc, q)) continue;      // not this pixel's winner
        if (!(sc > d.score[pix])) continue;              // strict improvement (:400, :490)
        const float4 pl = cplane[q];
        d.plane[pix] = pl;
        d.iplane[pix] = pl;
        d.depth[pix] = cdep[q];
        d.idepth[pix] = cdep[q];
        d.score[pix] = sc;
        d.iscore[pix] = sc;
        ++acc;
    }
    count_add(accepted, acc);
}

// ---------------------------------------------------------------- host side

void preinit_free(pmp_ctx* ctx) {
    PreState* s = ctx->pre;
    if (!s) return;
    void* b[] = {s->cpix, s->cplane, s->cdep, s->cscore, s->keys, s->xyz, s->accepted, s->prev_plane, s->prev_depth};
    for (void* x : b)
        if (x) (void)hipFree(x);
    delete s;
    ctx->pre = nullptr;
}

static int pre_state(pmp_ctx* ctx, PreState*& out) {
    if (!ctx->pre) {
        ctx->pre = new (std::nothrow) PreState();
        if (!ctx->pre) return set_error(ctx, PMP_ENOMEM, "pre-init state");
        PMP_HIP(ctx, hipMalloc(&ctx->pre->accepted, sizeof(unsigned long long)));
    }
    out = ctx->pre;
    return PMP_OK;
}

template <typename T>
static int grow(pmp_ctx* ctx, T*& p, size_t& cap, size_t n) {
    if (n <= cap) return PMP_OK;
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    PMP_HIP(ctx, hipMalloc(&p, n * sizeof(T)));
    cap = n;
    return PMP_OK;
}

// Buffers that share one capacity grow together; after a failed allocation
// the shared capacity is 0, so the next call reallocates all of them.
static int ensure_cands(pmp_ctx* ctx, PreState* s, size_t n, size_t pixels) {
    size_t c0 = s->cap, c1 = s->cap, c2 = s->cap, c3 = s->cap;
    int rc = grow(ctx, s->cpix, c0, n);
    if (!rc) rc = grow(ctx, s->cplane, c1, n);
    if (!rc) rc = grow(ctx, s->cdep, c2, n);
    if (!rc) rc = grow(ctx, s->cscore, c3, n);
    s->cap = rc ? 0 : c0;
    if (!rc) rc = grow(ctx, s->keys, s->keys_cap, pixels);
    return rc;
}

static int ensure_prev(pmp_ctx* ctx, PreState* s, size_t P) {
    size_t c0 = s->prev_cap, c1 = s->prev_cap;
    int rc = grow(ctx, s->prev_plane, c0, P);
    if (!rc) rc = grow(ctx, s->prev_depth, c1, P);
    s->prev_cap = rc ? 0 : c0;
    if (rc) s->prev_valid = false;
    return rc;
}

// Clears keys and counter, launches the candidate kernel and the apply
// kernel for the keyframe of descriptor sl (P pixels) on ctx->stream.
static int run_stage(pmp_ctx* ctx, QueueState* q, PreState* s, const SlotDesc* sl_dev, size_t P, const PreArgs& pa,
                     bool zero, bool warp, int64_t* accepted) {
    const SlotDesc PMP_CONST* sl = (const SlotDesc PMP_CONST*)sl_dev;
    int rc = q_fill_u32(ctx, ctx->stream, (unsigned int*)s->keys, 2 * P, 0u);
    if (!rc) rc = q_fill_u32(ctx, ctx->stream, (unsigned int*)s->accepted, 2, 0u);
    if (rc) return rc;
    const int cus = q->cus > 0 ? q->cus : 256;
    const unsigned grid = (unsigned)std::min<int64_t>((pa.n + 255) / 256, (int64_t)cus * 8);
    if (warp) {
        if (zero)
            hipLaunchKernelGGL(k_pre_warp<true>, dim3(grid), dim3(256), 0, ctx->stream, sl, pa, s->prev_plane,
                               s->prev_depth, s->cpix, s->cplane, s->cdep, s->cscore, s->keys);
        else
            hipLaunchKernelGGL(k_pre_warp<false>, dim3(grid), dim3(256), 0, ctx->stream, sl, pa, s->prev_plane,
                               s->prev_depth, s->cpix, s->cplane, s->cdep, s->cscore, s->keys);
    } else {
        if (zero)
            hipLaunchKernelGGL(k_pre_landmarks<true>, dim3(grid), dim3(256), 0, ctx->stream, sl, pa, s->xyz, s->cpix,
                               s->cplane, s->cdep, s->cscore, s->keys);
        else
            hipLaunchKernelGGL(k_pre_landmarks<false>, dim3(grid), dim3(256), 0, ctx->stream, sl, pa, s->xyz,
                               s->cpix, s->cplane, s->cdep, s->cscore, s->keys);
    }
    PMP_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_pre_apply, dim3(grid), dim3(256), 0, ctx->stream, sl, pa.n, s->cpix, s->cplane, s->cdep,
                       s->cscore, s->keys, s->accepted);
    PMP_HIP(ctx, hipGetLastError());
    if (accepted) {
        unsigned long long h = 0;
        PMP_HIP(ctx, hipMemcpyAsync(&h, s->accepted, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        *accepted = (int64_t)h;
    }
    return PMP_OK;
}

// Landmark seeding of one keyframe slot of a descriptor set (the served
// form): slot_dev is the slot's uploaded descriptor, whose maps hold the
// initial planes; xyz (3 n doubles) is pinned host memory that stays valid
// until the stream has passed this stage.  Stream-ordered on ctx->stream, no
// host synchronisation (scratch growth aside).
int preinit_landmarks_slot(pmp_ctx* ctx, const SlotDesc* slot_dev, int W, int H, const double* pose, int64_t n,
                           const double* xyz, bool zero) {
    if (n <= 0) return PMP_OK;
    if (n > 0x7FFFFFFF) return set_error(ctx, PMP_EINVAL, "too many landmarks");
    QueueState* q = nullptr;
    int rc = queue_state(ctx, q);
    PreState* s = nullptr;
    if (!rc) rc = pre_state(ctx, s);
    if (!rc) rc = ensure_cands(ctx, s, (size_t)n, (size_t)W * H);
    if (!rc) rc = grow(ctx, s->xyz, s->xyz_cap, 3 * (size_t)n);
    if (rc) return rc;
    PMP_HIP(ctx, hipMemcpyAsync(s->xyz, xyz, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    PreArgs pa;
    memcpy(pa.pose, pose, sizeof(pa.pose));
    memset(pa.prev_pose, 0, sizeof(pa.prev_pose));
    pa.n = (int)n;
    return run_stage(ctx, q, s, slot_dev, (size_t)W * H, pa, zero, false, nullptr);
}

int preinit_warp_slot(pmp_ctx* ctx, const SlotDesc* slot_dev, int W, int H, const double* pose,
                      const double* prev_pose, const float* prev_plane4, const float* prev_depth, bool zero) {
    const size_t P = (size_t)W * H;
    QueueState* q = nullptr;
    int rc = queue_state(ctx, q);
    PreState* s = nullptr;
    if (!rc) rc = pre_state(ctx, s);
    if (!rc) rc = ensure_cands(ctx, s, P, P);
    if (!rc) rc = ensure_prev(ctx, s, P);
    if (rc) return rc;
    PMP_HIP(ctx, hipMemcpyAsync(s->prev_plane, prev_plane4, P * sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
    PMP_HIP(ctx, hipMemcpyAsync(s->prev_depth, prev_depth, P * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    s->prev_valid = false;   // the buffers no longer hold a memorised one-keyframe result
    PreArgs pa;
    memcpy(pa.pose, pose, sizeof(pa.pose));
    memcpy(pa.prev_pose, prev_pose, sizeof(pa.prev_pose));
    pa.n = (int)P;
    return run_stage(ctx, q, s, slot_dev, P, pa, zero, true, nullptr);
}

}  // namespace pmp

using namespace pmp;

extern "C" {

int pmp_preinit_landmarks(pmp_ctx* ctx, const pmp_params* params, int64_t n, const double* xyz, const uint8_t* mask,
                          int64_t* accepted) {
    TraceRange trace_range("pmp_preinit_landmarks");
    if (!ctx || !params || n < 0 || (n > 0 && !xyz)) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (n > 0x7FFFFFFF) return set_error(ctx, PMP_EINVAL, "too many landmarks");
    if (accepted) *accepted = 0;
    if (!ctx->views_set) return set_error(ctx, PMP_ESTATE, "pmp_set_views has not been called");
    if (n == 0) return PMP_OK;
    const uint8_t* dmask = nullptr;
    int rc = stage_mask(ctx, mask, dmask);
    QueueState* q = nullptr;
    QueueRun r;
    if (!rc) rc = plain_prepare(ctx, params, dmask, true, 0, false, q, r);
    PreState* s = nullptr;
    if (!rc) rc = pre_state(ctx, s);
    if (!rc) rc = ensure_cands(ctx, s, (size_t)n, (size_t)ctx->W * ctx->H);
    if (!rc) rc = grow(ctx, s->xyz, s->xyz_cap, 3 * (size_t)n);
    if (rc) return rc;
    PMP_HIP(ctx, hipMemcpyAsync(s->xyz, xyz, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    PreArgs pa;
    memcpy(pa.pose, ctx->ref_pose, sizeof(pa.pose));
    memset(pa.prev_pose, 0, sizeof(pa.prev_pose));
    pa.n = (int)n;
    return run_stage(ctx, q, s, q->plain.d_slots, (size_t)ctx->W * ctx->H, pa, params->edge_mode == PMP_EDGE_ZERO,
                     false, accepted);
}

int pmp_memorize_result(pmp_ctx* ctx) {
    if (!ctx) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (!ctx->views_set) return set_error(ctx, PMP_ESTATE, "pmp_set_views has not been called");
    PreState* s = nullptr;
    int rc = pre_state(ctx, s);
    const size_t P = (size_t)ctx->W * ctx->H;
    if (!rc) rc = ensure_prev(ctx, s, P);
    if (rc) return rc;
    PMP_HIP(ctx, hipMemcpyAsync(s->prev_plane, ctx->ws.oplane, P * sizeof(float4), hipMemcpyDeviceToDevice,
                                ctx->stream));
    PMP_HIP(ctx, hipMemcpyAsync(s->prev_depth, ctx->ws.odepth, P * sizeof(float), hipMemcpyDeviceToDevice,
                                ctx->stream));
    memcpy(s->prev_pose, ctx->ref_pose, sizeof(s->prev_pose));
    s->prev_W = ctx->W;
    s->prev_H = ctx->H;
    s->prev_valid = true;
    return PMP_OK;
}

int pmp_forget_result(pmp_ctx* ctx) {
    if (!ctx) return PMP_EINVAL;
    if (ctx->pre) ctx->pre->prev_valid = false;
    return PMP_OK;
}

int pmp_preinit_warp(pmp_ctx* ctx, const pmp_params* params, const uint8_t* mask, int64_t* accepted) {
    TraceRange trace_range("pmp_preinit_warp");
    if (!ctx || !params) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (accepted) *accepted = 0;
    if (!ctx->views_set) return set_error(ctx, PMP_ESTATE, "pmp_set_views has not been called");
    PreState* s = ctx->pre;
    if (!s || !s->prev_valid) return PMP_OK;   // nothing memorised yet (depthmap.cc:437-439)
    if (s->prev_W != ctx->W || s->prev_H != ctx->H)
        return set_error(ctx, PMP_EINVAL, "the memorised result has another size");
    const uint8_t* dmask = nullptr;
    int rc = stage_mask(ctx, mask, dmask);
    QueueState* q = nullptr;
    QueueRun r;
    if (!rc) rc = plain_prepare(ctx, params, dmask, true, 0, false, q, r);
    const size_t P = (size_t)ctx->W * ctx->H;
    if (!rc) rc = ensure_cands(ctx, s, P, P);
    if (rc) return rc;
    PreArgs pa;
    memcpy(pa.pose, ctx->ref_pose, sizeof(pa.pose));
    memcpy(pa.prev_pose, s->prev_pose, sizeof(pa.prev_pose));
    pa.n = (int)P;
    return run_stage(ctx, q, s, q->plain.d_slots, P, pa, params->edge_mode == PMP_EDGE_ZERO, true, accepted);
}

}  // extern "C"
