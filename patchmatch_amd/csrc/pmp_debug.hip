// pmp_debug.hip -- diagnostics: device elementary functions and random normals
// evaluated on host-provided inputs, for bit-for-bit checks against the host.
#include <string.h>

#include <vector>

#include "pmp_internal.h"
#include "pmp_math.h"
#include "pmp_score.h"

namespace pmp {

__global__ void k_eval_math(int which, int n, const float* __restrict__ a, const float* __restrict__ b,
                            float* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float x = a[k];
    float r = 0.0f, s, c;
    switch (which) {
        case 0: r = rcp(x); break;
        case 1: r = rsqrt_pos(clampf(x, 1e-30f, 1e30f)); break;
        case 2: r = sqrt_(x); break;
        case 3: r = asin_(clampf(x, -1.0f, 1.0f)); break;
        case 4: r = acos_(clampf(x, -1.0f, 1.0f)); break;
        case 5: r = atan2_(x, b[k]); break;
        case 6: r = exp_(x); break;
        case 7: r = log_(x); break;
        case 8: sincos2pi(x, &s, &c); r = s; break;
        case 9: sincos2pi(x, &s, &c); r = c; break;
        case 10: lonlat(b[k], x, b[k], s, r); break;   // latitude of (b, a, b)
        case 11: lonlat(x, 0.0f, b[k], r, s); break;  // longitude of (a, 0, b)
        default: break;
    }
    out[k] = r;
}

__global__ void k_refine_normals(KfArgs a, int i, int j, int it, int n, float* __restrict__ out) {
    if (threadIdx.x != 0) return;
    float z[3 * kMaxRefine];
    refine_normals<3 * kMaxRefine>(a, i, j, it, n, z);
    for (int t = 0; t < n; ++t) out[t] = z[t];
}

__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

static int ensure_dbg(pmp_ctx* ctx, size_t floats) {
    if (ctx->dbg_cap >= floats) return PMP_OK;
    if (ctx->dbg) (void)hipFree(ctx->dbg);
    ctx->dbg = nullptr;
    ctx->dbg_cap = 0;
    PMP_HIP(ctx, hipMalloc(&ctx->dbg, floats * sizeof(float)));
    ctx->dbg_cap = floats;
    return PMP_OK;
}

}  // namespace pmp

using namespace pmp;

extern "C" {

int pmp_debug_eval_math(pmp_ctx* ctx, int32_t which, int64_t n, const float* a, const float* b, float* out) {
    if (!ctx || !a || !out || n < 0 || n > (1ll << 28) || which < 0 || which > 11) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (n == 0) return PMP_OK;
    int rc = ensure_dbg(ctx, 3 * (size_t)n);
    if (rc) return rc;
    float* da = ctx->dbg;
    float* db = ctx->dbg + n;
    float* dout = ctx->dbg + 2 * n;
    PMP_HIP(ctx, hipMemcpyAsync(da, a, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    PMP_HIP(ctx, hipMemcpyAsync(db, b ? b : a, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_eval_math, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, which, (int)n, da,
                       db, dout);
    PMP_HIP(ctx, hipGetLastError());
    PMP_HIP(ctx, hipMemcpyAsync(out, dout, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PMP_OK;
}

int pmp_debug_copy(pmp_ctx* ctx, int64_t bytes, int32_t repeats) {
    if (!ctx || bytes <= 0 || bytes % 16 != 0 || bytes > (1ll << 33) || repeats < 1) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    int rc = ensure_dbg(ctx, (size_t)bytes / 2);   // two halves of bytes each
    if (rc) return rc;
    const size_t n = (size_t)bytes / 16;
    const uint4* src = reinterpret_cast<const uint4*>(ctx->dbg);
    uint4* dst = reinterpret_cast<uint4*>(ctx->dbg) + n;
    for (int r = 0; r < repeats; ++r)
        hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, ctx->stream, src, dst, n);
    PMP_HIP(ctx, hipGetLastError());
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PMP_OK;
}

int pmp_debug_epochs(pmp_ctx* ctx, int32_t mode) {
    if (!ctx || mode < 0 || mode > 2) return PMP_EINVAL;
    ctx->epoch_mode = mode;
    return PMP_OK;
}

int pmp_debug_epoch_report(pmp_ctx* ctx, int32_t* tiles, int32_t* violations, int32_t* min_epoch,
                           int32_t* max_epoch) {
    if (!ctx || !tiles || !violations || !min_epoch || !max_epoch) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    if (!ctx->epochs || ctx->epoch_tiles.empty())
        return set_error(ctx, PMP_ESTATE, "no red-black pass has run with pass epochs on");
    std::vector<unsigned int> h(ctx->epoch_words);
    PMP_HIP(ctx, hipMemcpyAsync(h.data(), ctx->epochs, h.size() * sizeof(unsigned int), hipMemcpyDeviceToHost,
                                ctx->stream));
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // segments in slot order: the slot's tiles, then its violation count
    unsigned int lo = ~0u, hi = 0;
    long long n_tiles = 0, bad = 0;
    size_t off = 0;
    for (const int t : ctx->epoch_tiles) {
        for (int i = 0; i < t; ++i) {
            const unsigned int v = h[off + (size_t)i];
            lo = v < lo ? v : lo;
            hi = v > hi ? v : hi;
        }
        bad += h[off + (size_t)t];
        n_tiles += t;
        off += (size_t)t + 1;
    }
    *tiles = (int32_t)n_tiles;
    *violations = (int32_t)bad;
    *min_epoch = (int32_t)lo;
    *max_epoch = (int32_t)hi;
    return PMP_OK;
}

int pmp_debug_refine_normals(pmp_ctx* ctx, const pmp_params* params, int32_t width, int32_t i, int32_t j, int32_t it,
                             int32_t n, float* out) {
    if (!ctx || !params || !out || n < 0 || n > 3 * kMaxRefine || width < 1) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    int rc = ensure_dbg(ctx, 3 * kMaxRefine);
    if (rc) return rc;
    KfArgs a;
    memset(&a, 0, sizeof(a));
    a.W = width;
    a.key0 = (uint32_t)(params->seed & 0xFFFFFFFFu);
    a.key1 = (uint32_t)(params->seed >> 32);
    hipLaunchKernelGGL(k_refine_normals, dim3(1), dim3(64), 0, ctx->stream, a, i, j, it, n, ctx->dbg);
    PMP_HIP(ctx, hipGetLastError());
    PMP_HIP(ctx, hipMemcpyAsync(out, ctx->dbg, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PMP_OK;
}

}  // extern "C"
