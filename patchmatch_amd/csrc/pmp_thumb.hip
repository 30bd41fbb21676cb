// pmp_thumb.hip -- the frame thumbnails on the device (tracking_module.cc:82-93,
// data/frame.cc:53-54, :136-138): full-resolution camera images are
// converted to grey (cv::cvtColor) and resized to the depth-map size
// (cv::resize, INTER_LINEAR), the colour image and the mask resized too.
//
// 8-bit fixed point as OpenCV's published formulas:
//   grey   Y = (1868 B + 9617 G + 4899 R + 2^13) >> 14
//   resize source coordinate f = (d + 0.5) * (src/dst) - 0.5 (float), s =
//          floor(f), clamped to the first / last sample with f = 0; weights
//          round((1 - f) * 2048), round(f * 2048) (shorts); rows then
//          columns in int, (v + 2^21) >> 22.
// The tap tables are computed on the host; both kernels are HBM-bound.
#include <math.h>
#include <string.h>

#include <vector>

#include "pmp_internal.h"

namespace pmp {

struct Tap {
    int s0, s1;
    int a0, a1;
};

__global__ void k_to_gray(const uint8_t* __restrict__ src, int W, int H, int64_t pitch, int C, int order,
                          uint8_t* __restrict__ dst) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= W || i >= H) return;
    const uint8_t* p = src + (size_t)i * (size_t)pitch + (size_t)j * C;
    uint8_t y;
    if (C == 1 || order == PMP_COLOR_GRAY) {
        y = p[0];
    } else {
        const int r = order == PMP_COLOR_RGB ? p[0] : p[2], g = p[1], b = order == PMP_COLOR_RGB ? p[2] : p[0];
        y = (uint8_t)((b * 1868 + g * 9617 + r * 4899 + (1 << 13)) >> 14);
    }
    dst[(size_t)i * W + j] = y;
}

__global__ void k_resize(const uint8_t* __restrict__ src, int64_t pitch, int C, const Tap* __restrict__ xt,
                         const Tap* __restrict__ yt, uint8_t* __restrict__ dst, int dW, int dH, int64_t dpitch) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= dW || y >= dH) return;
    const Tap tx = xt[x], ty = yt[y];
    const uint8_t* r0 = src + (size_t)ty.s0 * (size_t)pitch;
    const uint8_t* r1 = src + (size_t)ty.s1 * (size_t)pitch;
    uint8_t* o = dst + (size_t)y * (size_t)dpitch + (size_t)x * C;
    for (int c = 0; c < C; ++c) {
        const int h0 = r0[tx.s0 * C + c] * tx.a0 + r0[tx.s1 * C + c] * tx.a1;
        const int h1 = r1[tx.s0 * C + c] * tx.a0 + r1[tx.s1 * C + c] * tx.a1;
        const int v = (h0 * ty.a0 + h1 * ty.a1 + (1 << 21)) >> 22;
        o[c] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
    }
}

static void taps(int src, int dst, std::vector<Tap>& t) {
    t.resize(dst);
    const double scale = (double)src / (double)dst;
    for (int d = 0; d < dst; ++d) {
        float f = (float)((d + 0.5) * scale - 0.5);
        int s = (int)floorf(f);
        f -= (float)s;
        if (s < 0) { f = 0.0f; s = 0; }
        if (s >= src - 1) { f = 0.0f; s = src - 1; }
        t[d].s0 = s;
        t[d].s1 = s + 1 < src ? s + 1 : src - 1;
        t[d].a0 = (int)(short)lrintf((1.0f - f) * 2048.0f);
        t[d].a1 = (int)(short)lrintf(f * 2048.0f);
    }
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

// Source bytes on the device (uploaded into the staging buffer when on host).
static int stage_src(pmp_ctx* ctx, const uint8_t* pixels, int W, int H, int64_t pitch, int C, bool on_device,
                     const uint8_t*& dsrc, int64_t& dpitch) {
    if (on_device) {
        dsrc = pixels;
        dpitch = pitch;
        return PMP_OK;
    }
    const size_t row = (size_t)W * C;
    int rc = grow(ctx, ctx->thumb.src, ctx->thumb.src_cap, row * H);
    if (rc) return rc;
    if ((rc = upload_host_rows(ctx, ctx->thumb.src, pixels, row, (size_t)H, (size_t)pitch, ctx->stream))) return rc;
    dsrc = ctx->thumb.src;
    dpitch = (int64_t)row;
    return PMP_OK;
}

static int launch_resize(pmp_ctx* ctx, const uint8_t* src, int W, int H, int64_t pitch, int C, uint8_t* dst, int dW,
                         int dH, int64_t dpitch) {
    std::vector<Tap> xt, yt;
    taps(W, dW, xt);
    taps(H, dH, yt);
    int rc = grow(ctx, ctx->thumb.taps, ctx->thumb.taps_cap, (size_t)(dW + dH) * sizeof(Tap));   // bytes
    if (rc) return rc;
    Tap* d_xt = (Tap*)ctx->thumb.taps;
    Tap* d_yt = d_xt + dW;
    PMP_HIP(ctx, hipMemcpyAsync(d_xt, xt.data(), dW * sizeof(Tap), hipMemcpyHostToDevice, ctx->stream));
    PMP_HIP(ctx, hipMemcpyAsync(d_yt, yt.data(), dH * sizeof(Tap), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_resize, dim3((dW + 255) / 256, dH), dim3(256), 0, ctx->stream, src, pitch, C, d_xt, d_yt, dst,
                       dW, dH, dpitch);
    PMP_HIP(ctx, hipGetLastError());
    // the host tap vectors are released on return: finish their copies first
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PMP_OK;
}

static int check_dims(pmp_ctx* ctx, int W, int H, int64_t pitch, int C, int dW, int dH) {
    if (W <= 0 || H <= 0 || dW <= 0 || dH <= 0) return set_error(ctx, PMP_EINVAL, "sizes must be positive");
    if (C != 1 && C != 3 && C != 4) return set_error(ctx, PMP_EINVAL, "channels must be 1, 3 or 4");
    if (pitch < (int64_t)W * C) return set_error(ctx, PMP_EINVAL, "pitch smaller than a row");
    return PMP_OK;
}

}  // namespace pmp

using namespace pmp;

extern "C" {

int pmp_resize(pmp_ctx* ctx, const uint8_t* src, int32_t width, int32_t height, int64_t pitch, int32_t channels,
               uint8_t* dst, int32_t out_width, int32_t out_height, int32_t on_device) {
    if (!ctx || !src || !dst) return PMP_EINVAL;
    int rc = check_dims(ctx, width, height, pitch, channels, out_width, out_height);
    if (rc) return rc;
    PMP_ON_DEVICE(ctx);
    const uint8_t* dsrc;
    int64_t dpitch;
    if ((rc = stage_src(ctx, src, width, height, pitch, channels, on_device != 0, dsrc, dpitch))) return rc;
    uint8_t* ddst = dst;
    const size_t out_bytes = (size_t)out_width * out_height * channels;
    if (!on_device) {
        if ((rc = grow(ctx, ctx->thumb.dst, ctx->thumb.dst_cap, out_bytes))) return rc;
        ddst = ctx->thumb.dst;
    }
    if ((rc = launch_resize(ctx, dsrc, width, height, dpitch, channels, ddst, out_width, out_height,
                            (int64_t)out_width * channels)))
        return rc;
    if (!on_device) {
        PMP_HIP(ctx, hipMemcpyAsync(dst, ddst, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
        PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return PMP_OK;
}

int pmp_image_put_thumbnail(pmp_ctx* ctx, uint64_t id, const uint8_t* pixels, int32_t width, int32_t height,
                            int64_t pitch, int32_t channels, int32_t color_order, int32_t out_width,
                            int32_t out_height, int32_t on_device) {
    if (!ctx || !pixels) return PMP_EINVAL;
    int rc = check_dims(ctx, width, height, pitch, channels, out_width, out_height);
    if (rc) return rc;
    if (color_order < PMP_COLOR_GRAY || color_order > PMP_COLOR_BGR)
        return set_error(ctx, PMP_EINVAL, "unknown color order");
    PMP_ON_DEVICE(ctx);
    const uint8_t* dsrc;
    int64_t dpitch;
    if ((rc = stage_src(ctx, pixels, width, height, pitch, channels, on_device != 0, dsrc, dpitch))) return rc;
    // grey at full resolution first (tracking_module.cc:82-83), then the resize (frame.cc:54)
    const uint8_t* grey = dsrc;
    int64_t gpitch = dpitch;
    if (channels != 1) {
        if ((rc = grow(ctx, ctx->thumb.gray, ctx->thumb.gray_cap, (size_t)width * height))) return rc;
        hipLaunchKernelGGL(k_to_gray, dim3((width + 255) / 256, height), dim3(256), 0, ctx->stream, dsrc, width,
                           height, dpitch, channels, color_order, ctx->thumb.gray);
        PMP_HIP(ctx, hipGetLastError());
        grey = ctx->thumb.gray;
        gpitch = width;
    }
    // into the pool under id (replacing it), like pmp_image_put
    auto it = ctx->images.find(id);
    if (it != ctx->images.end()) {
        PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (it->second.gray) (void)hipFree(it->second.gray);
        if (it->second.quads) (void)hipFree(it->second.quads);
        ctx->images.erase(it);
    }
    Image im;
    im.W = out_width;
    im.H = out_height;
    im.bytes = (size_t)out_width * out_height;
    PMP_HIP(ctx, hipMalloc(&im.gray, im.bytes));
    rc = launch_resize(ctx, grey, width, height, gpitch, 1, im.gray, out_width, out_height, out_width);
    if (rc) {
        (void)hipFree(im.gray);
        return rc;
    }
    ctx->images[id] = im;
    return PMP_OK;
}

}  // extern "C"
