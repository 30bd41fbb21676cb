// pmp_fusion.hip -- geometric-consistency filter (Cleaner) and point fusion
// (Pruner) kernels and their C-ABI.  Both are streaming, HBM-bound passes.
//
// Geometry is in double as in the reference (Backproject_equirect
// depthmap.cc:106-111, ReprojectEquirect :143-161); angles and pixel
// coordinates use the same float functions as the score kernels.
//
// The Pruner (depthmap.cc:1535-1696) visits head pixels in raster order and
// may zero pixels of later maps, so its result depends on that order.  The
// kernels here reproduce the sequential result exactly with one phase per
// later map o (a head pixel leaves the loop at the first map that drops it,
// so its state before map o depends only on phases < o):
//   mark   every live head pixel whose reprojection into map o is consistent
//          and nearer (depth < map value) offers its raster index to the
//          target pixel with atomicMin: the minimum is the pixel that zeroes
//          the target in the sequential loop;
//   drop   a live, consistent head pixel with a raster index below the
//          target's zeroing index meets the unmodified value with depth >= it
//          and is dropped; pixels after the zeroing one read 0 and continue;
//   apply  targets with a zeroing index are set to 0 (persisting into later
//          prunes, as in the reference).
#include <math.h>
#include <string.h>

#include <climits>
#include <mutex>
#include <vector>

#include "pmp_trace.h"
#include "pmp_internal.h"
#include "pmp_geo.h"
#include "pmp_math.h"

namespace pmp {

constexpr int kMaxMaps = 16;

struct FuseArgs {
    GeoArgs g;
    int n;
    float thr;
    double poses[kMaxMaps][12];
};

// ---------------------------------------------------------------- Cleaner
// Optional per-view tests on top of the reference's depth test (pmp_clean_ex;
// the reference has neither).  Both are off in pmp_clean.
struct CleanChecks {
    const float* planes;  // n maps of W*H*4 floats (camera-frame p = n/d), or null
    float px2;            // squared pixel bound of the round trip; <= 0: off
    float iw, ih;         // 1/W, 1/H
    double cos2;          // squared cosine of the normal-angle bound; <= 0: off
};

// Round trip: the point that map o holds at (x, y), depth dr, projected back
// into map 0 lands within sqrt(px2) pixels of (i, j) (longitude wrapped).
__device__ __forceinline__ bool round_trip_ok(const FuseArgs& a, const CleanChecks& c, int o, int i, int j, float x,
                                              float y, float dr) {
    float bx, by, bz;
    bearing_at(x, y, c.iw, c.ih, bx, by, bz);
    double X0, X1, X2;
    backproject_dir(a.poses[o], (double)bx, (double)by, (double)bz, dr, X0, X1, X2);
    float x0, y0, d0;
    if (reproject(a.g, a.poses[0], X0, X1, X2, x0, y0, d0)) return false;
    const float Wf = (float)a.g.W;
    float dx = x0 - (float)j;
    if (dx > 0.5f * Wf) dx = dx - Wf;
    else if (dx < -0.5f * Wf) dx = dx + Wf;
    const float dy = y0 - (float)i;
    return fma_(dx, dx, dy * dy) <= c.px2;
}

// Normals: the world-frame planes R^T p of map 0 at (i, j) and of map o at
// the pixel nearest (x, y) are at most acos(sqrt(cos2)) apart.
__device__ __forceinline__ bool normal_ok(const FuseArgs& a, const CleanChecks& c, int o, size_t k, float x, float y) {
    const int W = a.g.W, H = a.g.H;
    const int ix = min((int)(x + 0.5f), W - 1), iy = min((int)(y + 0.5f), H - 1);
    const size_t P = (size_t)W * H;
    const float4 p = reinterpret_cast<const float4*>(c.planes)[k];
    const float4 q = reinterpret_cast<const float4*>(c.planes)[(size_t)o * P + (size_t)iy * W + ix];
    const double* R = a.poses[0];
    const double* S = a.poses[o];
    const double p0 = (double)p.x, p1 = (double)p.y, p2 = (double)p.z;
    const double q0 = (double)q.x, q1 = (double)q.y, q2 = (double)q.z;
    const double a0 = R[0] * p0 + R[3] * p1 + R[6] * p2;
    const double a1 = R[1] * p0 + R[4] * p1 + R[7] * p2;
    const double a2 = R[2] * p0 + R[5] * p1 + R[8] * p2;
    const double b0 = S[0] * q0 + S[3] * q1 + S[6] * q2;
    const double b1 = S[1] * q0 + S[4] * q1 + S[7] * q2;
    const double b2 = S[2] * q0 + S[5] * q1 + S[8] * q2;
    const double d = a0 * b0 + a1 * b1 + a2 * b2;
    if (!(d > 0.0)) return false;
    const double na = a0 * a0 + a1 * a1 + a2 * a2, nb = b0 * b0 + b1 * b1 + b2 * b2;
    return d * d >= c.cos2 * (na * nb);
}

__global__ void k_clean(FuseArgs a, CleanChecks c, const float* __restrict__ maps, int min_consistent,
                        float* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    const int W = a.g.W, H = a.g.H;
    if (j >= W) return;
    const size_t P = (size_t)W * H;
    const size_t k = (size_t)i * W + j;
    const float dep = maps[k];
    float res = 0.0f;
    if (!((double)dep < 1e-6)) {
        double X0, X1, X2;
        backproject(a.g, a.poses[0], i, j, dep, X0, X1, X2);
        int consistent = 1;
        for (int o = 1; o < a.n; ++o) {
            float x, y, dp;
            if (reproject(a.g, a.poses[o], X0, X1, X2, x, y, dp)) continue;
            const int iy = (int)y, ix = (int)x;
            if (!(iy >= 0 && iy < H && ix >= 0 && ix < W)) continue;
            const float dr = lin_interp(maps + (size_t)o * P, W, H, y, x);
            if (!(abs_(dr - dp) < dp * a.thr)) continue;
            if (c.px2 > 0.0f && !round_trip_ok(a, c, o, i, j, x, y, dr)) continue;
            if (c.cos2 > 0.0 && !normal_ok(a, c, o, k, x, y)) continue;
            consistent++;
        }
        if (consistent >= min_consistent) res = dep;
    }
    out[k] = res;
}

// ---------------------------------------------------------------- Pruner
struct PruneMaps {
    const float* head;
    float* other[kMaxMaps];
};

// depthmap.cc:1559-1580 (the i-1>0 / j-1>0 tests skip row 0 / column 0 neighbours)
__global__ void k_prune_init(FuseArgs a, PruneMaps m, uint8_t* __restrict__ alive, float* __restrict__ pruned) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    const int W = a.g.W, H = a.g.H;
    if (j >= W) return;
    const size_t k = (size_t)i * W + j;
    const float* dm = m.head;
    const float dep = dm[k];
    bool filt = dep == 0.0f;
    if (i - 1 > 0 && i - 1 < H) filt = filt || dm[k - W] == 0.0f;
    if (j - 1 > 0 && j - 1 < W) filt = filt || dm[k - 1] == 0.0f;
    if (i + 1 > 0 && i + 1 < H) filt = filt || dm[k + W] == 0.0f;
    if (j + 1 > 0 && j + 1 < W) filt = filt || dm[k + 1] == 0.0f;
    alive[k] = filt ? 0 : 1;
    pruned[k] = filt ? 0.0f : dep;
}

// Target pixel in map o and the outcome of the reference's tests
// (depthmap.cc:1609-1625): returns false when the loop continues without
// touching map o; otherwise q = target index and nearer = depth < value.
__device__ __forceinline__ bool prune_hit(const FuseArgs& a, const PruneMaps& m, int o, int i, int j, float dep,
                                          int& q, bool& nearer) {
    const int W = a.g.W, H = a.g.H;
    double X0, X1, X2;
    backproject(a.g, a.poses[0], i, j, dep, X0, X1, X2);
    float x, y, dp;
    if (reproject(a.g, a.poses[o], X0, X1, X2, x, y, dp)) return false;
    const int iu = (int)x, iv = (int)y;
    if (!(iv >= 0 && iv < H && iu >= 0 && iu < W) || (double)dp < 1e-5) return false;
    q = iv * W + iu;
    const float dr = m.other[o][q];
    if (dr == 0.0f || abs_(dr - dp) > a.thr * dp) return false;
    nearer = dep < dr;
    return true;
}

__global__ void k_prune_mark(FuseArgs a, PruneMaps m, int o, const uint8_t* __restrict__ alive,
                             int32_t* __restrict__ zmin) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= a.g.W) return;
    const int k = i * a.g.W + j;
    if (!alive[k]) return;
    int q;
    bool nearer;
    if (prune_hit(a, m, o, i, j, m.head[k], q, nearer) && nearer) atomicMin(&zmin[q], k);
}

__global__ void k_prune_drop(FuseArgs a, PruneMaps m, int o, uint8_t* __restrict__ alive,
                             const int32_t* __restrict__ zmin, float* __restrict__ pruned) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= a.g.W) return;
    const int k = i * a.g.W + j;
    if (!alive[k]) return;
    int q;
    bool nearer;
    if (prune_hit(a, m, o, i, j, m.head[k], q, nearer) && k < zmin[q]) {
        alive[k] = 0;
        pruned[k] = 0.0f;
    }
}

__global__ void k_prune_apply(float* __restrict__ other, int32_t* __restrict__ zmin, int n) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    if (zmin[q] != INT_MAX) {
        other[q] = 0.0f;
        zmin[q] = INT_MAX;
    }
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int32_t v, int n) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = v;
}

constexpr int kScanBlock = 256;

__device__ __forceinline__ int block_prefix(bool flag, int* s_wave, int& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long bal = __ballot(flag);
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) s_wave[wave] = __popcll(bal);
    __syncthreads();
    int off = 0;
    total = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) {
        if (w < wave) off += s_wave[w];
        total += s_wave[w];
    }
    return off + before;
}

__global__ __launch_bounds__(kScanBlock) void k_prune_count(const uint8_t* __restrict__ alive, int n,
                                                            int32_t* __restrict__ blk) {
    __shared__ int s_wave[kScanBlock / 64];
    const int k = blockIdx.x * kScanBlock + threadIdx.x;
    int total;
    block_prefix(k < n && alive[k], s_wave, total);
    if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void k_prune_emit(FuseArgs a, const float* __restrict__ head,
                                                           const uint8_t* __restrict__ bgr,
                                                           const uint8_t* __restrict__ alive,
                                                           const int32_t* __restrict__ blk_off,
                                                           double* __restrict__ pts, uint8_t* __restrict__ cols,
                                                           int32_t* __restrict__ kps) {
    __shared__ int s_wave[kScanBlock / 64];
    const int W = a.g.W;
    const int n = W * a.g.H;
    const int k = blockIdx.x * kScanBlock + threadIdx.x;
    const bool f = k < n && alive[k];
    int total;
    const int pos = blk_off[blockIdx.x] + block_prefix(f, s_wave, total);
    if (!f) return;
    const int i = k / W, j = k - i * W;
    double X0, X1, X2;
    backproject(a.g, a.poses[0], i, j, head[k], X0, X1, X2);
    pts[3 * (size_t)pos] = X0;
    pts[3 * (size_t)pos + 1] = X1;
    pts[3 * (size_t)pos + 2] = X2;
    cols[3 * (size_t)pos] = bgr[3 * (size_t)k];
    cols[3 * (size_t)pos + 1] = bgr[3 * (size_t)k + 1];
    cols[3 * (size_t)pos + 2] = bgr[3 * (size_t)k + 2];
    kps[2 * (size_t)pos] = i;
    kps[2 * (size_t)pos + 1] = j;
}

// ---- the Cleaner without allocation (pmp_clean_noalloc)
// Its device memory is one file-scope array per GPU: the staged depth maps,
// the cleaned map and the two bearing tables of the call.  In host code the
// array's name is only a stand-in: the device address comes from
// hipGetSymbolAddress, looked up once per device and kept; only that address
// reaches kernels and copies.  Every call on one GPU uses the same array, so
// calls that overlap on one GPU race.
constexpr size_t kCleanScratchFloats = size_t(1) << 22;   // 16 MiB per GPU
static __device__ __align__(256) float scratch[kCleanScratchFloats];

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static float* g_scratch[kScratchDevices] = {};

// The device address of this GPU's scratch array (ctx's device current).
static int clean_scratch(pmp_ctx* ctx, float*& out) {
    if (ctx->device < 0 || ctx->device >= kScratchDevices)
        return set_error(ctx, PMP_EINVAL, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[ctx->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return check_hip(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                    "hipGetSymbolAddress(scratch)");
        g_scratch[ctx->device] = (float*)p;
    }
    out = g_scratch[ctx->device];
    return PMP_OK;
}

// Writes the W + H bearing-table entries of a W x H image to the device
// tables, through a fixed host buffer (no allocation).
static int write_tables(pmp_ctx* ctx, int W, int H, float2* rowtab, float2* coltab) {
    constexpr int kChunk = 512;
    float2 buf[kChunk];
    for (int side = 0; side < 2; ++side) {
        const int n = side == 0 ? H : W;
        float2* dst = side == 0 ? rowtab : coltab;
        for (int k0 = 0; k0 < n; k0 += kChunk) {
            const int m = n - k0 < kChunk ? n - k0 : kChunk;
            for (int k = 0; k < m; ++k) buf[k] = side == 0 ? table_row(k0 + k, H) : table_col(k0 + k, W);
            PMP_HIP(ctx, hipMemcpyAsync(dst + k0, buf, m * sizeof(float2), hipMemcpyHostToDevice, ctx->stream));
            PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));   // buf is rewritten next
        }
    }
    return PMP_OK;
}

static int geo_args(pmp_ctx* ctx, int W, int H, GeoArgs& g) {
    Tables tb;
    int rc = get_tables(ctx, W, H, ctx->last_sigma, tb);
    if (rc) return rc;
    g.W = W; g.H = H;
    g.rowtab = tb.row; g.coltab = tb.col;
    g.ku = (float)((double)W / (2.0 * M_PI));
    g.hu = (float)(0.5 * W);
    g.kv = (float)((double)H / M_PI);
    g.hv = (float)(0.5 * H);
    return PMP_OK;
}

}  // namespace pmp

using namespace pmp;

extern "C" {

static int clean_impl(pmp_ctx* ctx, int32_t width, int32_t height, int32_t n_maps, const float* const* depth_maps,
                      const float* const* plane4_maps, const pmp_pose* poses, float same_depth_thr,
                      int32_t min_consistent, float max_reproj_px, float max_normal_deg, float* cleaned) {
    if (!ctx || !depth_maps || !poses || !cleaned || width < 2 || height < 2) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (n_maps < 1 || n_maps > kMaxMaps) return set_error(ctx, PMP_EINVAL, "n_maps must be in [1, 16]");
    if (isnan(max_reproj_px) || !(max_normal_deg < 90.0f))
        return set_error(ctx, PMP_EINVAL, "max_reproj_px must not be NaN; max_normal_deg must be below 90");
    const bool use_normals = max_normal_deg > 0.0f;
    if (use_normals && !plane4_maps) return set_error(ctx, PMP_EINVAL, "the normal test needs the plane maps");
    FuseArgs a;
    memset(&a, 0, sizeof(a));
    int rc = geo_args(ctx, width, height, a.g);
    if (rc) return rc;
    a.n = n_maps;
    a.thr = same_depth_thr;
    for (int o = 0; o < n_maps; ++o) {
        memcpy(a.poses[o], poses[o].R, 9 * sizeof(double));
        memcpy(a.poses[o] + 9, poses[o].t, 3 * sizeof(double));
    }
    const size_t P = (size_t)width * height;
    const size_t poff = ((n_maps + 1) * P + 3) / 4 * 4;   // planes start 16-byte aligned
    const size_t need = use_normals ? poff + 4 * (size_t)n_maps * P : (n_maps + 1) * P;
    if (ctx->clean_cap < need) {
        if (ctx->clean_maps) (void)hipFree(ctx->clean_maps);
        ctx->clean_maps = nullptr;
        ctx->clean_cap = 0;
        PMP_HIP(ctx, hipMalloc(&ctx->clean_maps, need * sizeof(float)));
        ctx->clean_cap = need;
    }
    for (int o = 0; o < n_maps; ++o) {
        if (!depth_maps[o]) return set_error(ctx, PMP_EINVAL, "null depth map");
        PMP_HIP(ctx, hipMemcpyAsync(ctx->clean_maps + o * P, depth_maps[o], P * sizeof(float), hipMemcpyHostToDevice,
                                    ctx->stream));
    }
    float* dout = ctx->clean_maps + n_maps * P;
    CleanChecks c;
    c.planes = nullptr;
    c.px2 = max_reproj_px > 0.0f ? max_reproj_px * max_reproj_px : 0.0f;
    c.iw = (float)(1.0 / width);
    c.ih = (float)(1.0 / height);
    c.cos2 = 0.0;
    if (use_normals) {
        float* planes = ctx->clean_maps + poff;
        for (int o = 0; o < n_maps; ++o) {
            if (!plane4_maps[o]) return set_error(ctx, PMP_EINVAL, "null plane map");
            PMP_HIP(ctx, hipMemcpyAsync(planes + 4 * o * P, plane4_maps[o], 4 * P * sizeof(float),
                                        hipMemcpyHostToDevice, ctx->stream));
        }
        c.planes = planes;
        const double cs = cos((double)max_normal_deg * (M_PI / 180.0));
        c.cos2 = cs * cs;
    }
    dim3 blk(256), grd((width + 255) / 256, height);
    hipLaunchKernelGGL(k_clean, grd, blk, 0, ctx->stream, a, c, ctx->clean_maps, min_consistent, dout);
    PMP_HIP(ctx, hipGetLastError());
    PMP_HIP(ctx, hipMemcpyAsync(cleaned, dout, P * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PMP_OK;
}

int pmp_clean(pmp_ctx* ctx, int32_t width, int32_t height, int32_t n_maps, const float* const* depth_maps,
              const pmp_pose* poses, float same_depth_thr, int32_t min_consistent, float* cleaned) {
    TraceRange trace_range("pmp_clean");
    return clean_impl(ctx, width, height, n_maps, depth_maps, nullptr, poses, same_depth_thr, min_consistent, 0.0f,
                      0.0f, cleaned);
}

int pmp_clean_ex(pmp_ctx* ctx, int32_t width, int32_t height, int32_t n_maps, const float* const* depth_maps,
                 const float* const* plane4_maps, const pmp_pose* poses, float same_depth_thr, int32_t min_consistent,
                 float max_reproj_px, float max_normal_deg, float* cleaned) {
    TraceRange trace_range("pmp_clean_ex");
    return clean_impl(ctx, width, height, n_maps, depth_maps, plane4_maps, poses, same_depth_thr, min_consistent,
                      max_reproj_px, max_normal_deg, cleaned);
}

int pmp_clean_noalloc(pmp_ctx* ctx, int32_t width, int32_t height, int32_t n_maps, const float* const* depth_maps,
                      const pmp_pose* poses, float same_depth_thr, int32_t min_consistent, float* cleaned) {
    if (!ctx || !depth_maps || !poses || !cleaned || width < 2 || height < 2) return PMP_EINVAL;
    if (n_maps < 1 || n_maps > kMaxMaps) return set_error(ctx, PMP_EINVAL, "n_maps must be in [1, 16]");
    const size_t P = (size_t)width * height;
    const size_t tab_off = ((size_t)(n_maps + 1) * P + 1) / 2 * 2;   // float2 tables start 8-byte aligned
    const size_t need = tab_off + 2 * ((size_t)width + (size_t)height);
    if (need > kCleanScratchFloats)
        return set_error(ctx, PMP_EINVAL,
                         "pmp_clean_noalloc holds at most 4194304 floats ((n_maps + 1) * width * height + "
                         "2 * (width + height)); use pmp_clean");
    for (int o = 0; o < n_maps; ++o)
        if (!depth_maps[o]) return set_error(ctx, PMP_EINVAL, "null depth map");
    PMP_ON_DEVICE(ctx);
    float* base = nullptr;
    int rc = clean_scratch(ctx, base);
    if (rc) return rc;
    float* dout = base + (size_t)n_maps * P;
    float2* rowtab = (float2*)(base + tab_off);
    float2* coltab = rowtab + height;
    rc = write_tables(ctx, width, height, rowtab, coltab);
    if (rc) return rc;
    FuseArgs a;
    memset(&a, 0, sizeof(a));
    a.g.W = width;
    a.g.H = height;
    a.g.rowtab = rowtab;
    a.g.coltab = coltab;
    a.g.ku = (float)((double)width / (2.0 * M_PI));
    a.g.hu = (float)(0.5 * width);
    a.g.kv = (float)((double)height / M_PI);
    a.g.hv = (float)(0.5 * height);
    a.n = n_maps;
    a.thr = same_depth_thr;
    for (int o = 0; o < n_maps; ++o) {
        memcpy(a.poses[o], poses[o].R, 9 * sizeof(double));
        memcpy(a.poses[o] + 9, poses[o].t, 3 * sizeof(double));
        PMP_HIP(ctx, hipMemcpyAsync(base + (size_t)o * P, depth_maps[o], P * sizeof(float), hipMemcpyHostToDevice,
                                    ctx->stream));
    }
    CleanChecks c;
    c.planes = nullptr;
    c.px2 = 0.0f;
    c.iw = (float)(1.0 / width);
    c.ih = (float)(1.0 / height);
    c.cos2 = 0.0;
    dim3 blk(256), grd((width + 255) / 256, height);
    hipLaunchKernelGGL(k_clean, grd, blk, 0, ctx->stream, a, c, (const float*)base, min_consistent, dout);
    PMP_HIP(ctx, hipGetLastError());
    PMP_HIP(ctx, hipMemcpyAsync(cleaned, dout, P * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PMP_OK;
}

int pmp_pruner_add_view(pmp_ctx* ctx, const float* cleaned_depth, const pmp_pose* pose, const uint8_t* bgr,
                        int32_t width, int32_t height) {
    if (!ctx || !cleaned_depth || !pose || !bgr || width < 2 || height < 2) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (!ctx->pruner.empty() && (ctx->pruner[0].W != width || ctx->pruner[0].H != height))
        return set_error(ctx, PMP_EINVAL, "all pruner views must have the same size");
    if ((int)ctx->pruner.size() >= kMaxMaps) return set_error(ctx, PMP_EINVAL, "pruner queue is full (16)");
    PrunerView v;
    v.W = width;
    v.H = height;
    const size_t P = (size_t)width * height;
    PMP_HIP(ctx, hipMalloc(&v.depth, P * sizeof(float)));
    hipError_t e = hipMalloc(&v.bgr, 3 * P);
    if (e != hipSuccess) {
        (void)hipFree(v.depth);
        return check_hip(ctx, e, "hipMalloc bgr");
    }
    memcpy(v.pose, pose->R, 9 * sizeof(double));
    memcpy(v.pose + 9, pose->t, 3 * sizeof(double));
    PMP_HIP(ctx, hipMemcpyAsync(v.depth, cleaned_depth, P * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    PMP_HIP(ctx, hipMemcpyAsync(v.bgr, bgr, 3 * P, hipMemcpyHostToDevice, ctx->stream));
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->pruner.push_back(v);
    return PMP_OK;
}

int pmp_pruner_prune(pmp_ctx* ctx, float same_depth_thr, float* pruned_depth, int64_t cap, double* pts, uint8_t* cols,
                     int32_t* kps, int64_t* count) {
    TraceRange trace_range("pmp_pruner_prune");
    if (!ctx || cap < 0 || (cap > 0 && (!pts || !cols || !kps))) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (ctx->pruner.empty()) {
        if (count) *count = 0;
        return PMP_OK;
    }
    const int W = ctx->pruner[0].W, H = ctx->pruner[0].H;
    const int n = (int)ctx->pruner.size();
    const size_t P = (size_t)W * H;
    FuseArgs a;
    memset(&a, 0, sizeof(a));
    int rc = geo_args(ctx, W, H, a.g);
    if (rc) return rc;
    a.n = n;
    a.thr = same_depth_thr;
    PruneMaps m;
    memset(&m, 0, sizeof(m));
    m.head = ctx->pruner[0].depth;
    for (int o = 0; o < n; ++o) {
        memcpy(a.poses[o], ctx->pruner[o].pose, 12 * sizeof(double));
        m.other[o] = ctx->pruner[o].depth;
    }
    const int nblk = (int)((P + kScanBlock - 1) / kScanBlock);
    if (ctx->prune_cap < P) {
        void* olds[] = {ctx->prune_z, ctx->prune_alive, ctx->prune_out, ctx->prune_pts, ctx->prune_cols,
                        ctx->prune_kps, ctx->prune_blk};
        for (void* b : olds)
            if (b) (void)hipFree(b);
        ctx->prune_z = nullptr; ctx->prune_alive = nullptr; ctx->prune_out = nullptr; ctx->prune_pts = nullptr;
        ctx->prune_cols = nullptr; ctx->prune_kps = nullptr; ctx->prune_blk = nullptr;
        ctx->prune_cap = 0;
        PMP_HIP(ctx, hipMalloc(&ctx->prune_z, P * sizeof(int32_t)));
        PMP_HIP(ctx, hipMalloc(&ctx->prune_alive, P));
        PMP_HIP(ctx, hipMalloc(&ctx->prune_out, P * sizeof(float)));
        PMP_HIP(ctx, hipMalloc(&ctx->prune_pts, 3 * P * sizeof(double)));
        PMP_HIP(ctx, hipMalloc(&ctx->prune_cols, 3 * P));
        PMP_HIP(ctx, hipMalloc(&ctx->prune_kps, 2 * P * sizeof(int32_t)));
        PMP_HIP(ctx, hipMalloc(&ctx->prune_blk, ((P + kScanBlock - 1) / kScanBlock) * sizeof(int32_t)));
        ctx->prune_cap = P;
    }
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk), dim3(kScanBlock), 0, ctx->stream, ctx->prune_z, INT_MAX, (int)P);
    dim3 blk(256), grd((W + 255) / 256, H);
    hipLaunchKernelGGL(k_prune_init, grd, blk, 0, ctx->stream, a, m, ctx->prune_alive, ctx->prune_out);
    for (int o = 1; o < n; ++o) {
        hipLaunchKernelGGL(k_prune_mark, grd, blk, 0, ctx->stream, a, m, o, ctx->prune_alive, ctx->prune_z);
        hipLaunchKernelGGL(k_prune_drop, grd, blk, 0, ctx->stream, a, m, o, ctx->prune_alive, ctx->prune_z,
                           ctx->prune_out);
        hipLaunchKernelGGL(k_prune_apply, dim3(nblk), dim3(kScanBlock), 0, ctx->stream, m.other[o], ctx->prune_z,
                           (int)P);
    }
    hipLaunchKernelGGL(k_prune_count, dim3(nblk), dim3(kScanBlock), 0, ctx->stream, ctx->prune_alive, (int)P,
                       ctx->prune_blk);
    PMP_HIP(ctx, hipGetLastError());
    std::vector<int32_t> hb(nblk);
    PMP_HIP(ctx, hipMemcpyAsync(hb.data(), ctx->prune_blk, nblk * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int64_t total = 0;
    for (int b = 0; b < nblk; ++b) {
        int32_t c = hb[b];
        hb[b] = (int32_t)total;
        total += c;
    }
    PMP_HIP(ctx, hipMemcpyAsync(ctx->prune_blk, hb.data(), nblk * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_prune_emit, dim3(nblk), dim3(kScanBlock), 0, ctx->stream, a, m.head, ctx->pruner[0].bgr,
                       ctx->prune_alive, ctx->prune_blk, ctx->prune_pts, ctx->prune_cols, ctx->prune_kps);
    PMP_HIP(ctx, hipGetLastError());
    const int64_t nc = total < cap ? total : cap;
    if (nc > 0) {
        PMP_HIP(ctx, hipMemcpyAsync(pts, ctx->prune_pts, nc * 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        PMP_HIP(ctx, hipMemcpyAsync(cols, ctx->prune_cols, nc * 3, hipMemcpyDeviceToHost, ctx->stream));
        PMP_HIP(ctx, hipMemcpyAsync(kps, ctx->prune_kps, nc * 2 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    }
    if (pruned_depth)
        PMP_HIP(ctx, hipMemcpyAsync(pruned_depth, ctx->prune_out, P * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (count) *count = total;
    return PMP_OK;
}

int pmp_pruner_pop_head(pmp_ctx* ctx) {
    if (!ctx) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (ctx->pruner.empty()) return PMP_OK;
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    (void)hipFree(ctx->pruner[0].depth);
    (void)hipFree(ctx->pruner[0].bgr);
    ctx->pruner.erase(ctx->pruner.begin());
    return PMP_OK;
}

int pmp_pruner_reset(pmp_ctx* ctx) {
    if (!ctx) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto& v : ctx->pruner) {
        (void)hipFree(v.depth);
        (void)hipFree(v.bgr);
    }
    ctx->pruner.clear();
    return PMP_OK;
}

int pmp_pruner_size(const pmp_ctx* ctx) { return ctx ? (int)ctx->pruner.size() : PMP_EINVAL; }

int pmp_pruner_get_map(pmp_ctx* ctx, int32_t index, float* depth) {
    if (!ctx || !depth) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (index < 0 || index >= (int)ctx->pruner.size()) return set_error(ctx, PMP_EINVAL, "pruner index out of range");
    const PrunerView& v = ctx->pruner[index];
    PMP_HIP(ctx, hipMemcpyAsync(depth, v.depth, (size_t)v.W * v.H * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PMP_OK;
}

}  // extern "C"
