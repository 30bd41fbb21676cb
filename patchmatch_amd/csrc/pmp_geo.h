// pmp_geo.h -- device geometry of the Cleaner, the Pruner and the
// pre-initialisation: double-precision rigid transforms as in the reference,
// angles and pixel coordinates with this project's float functions.
#pragma once
#include "pmp_internal.h"
#include "pmp_math.h"

namespace pmp {

struct GeoArgs {
    int W, H;
    const float2* rowtab;
    const float2* coltab;
    float ku, hu, kv, hv;
};

// R^T (d b - t) for a bearing b
__device__ __forceinline__ void backproject_dir(const double* pose, double bx, double by, double bz, float dep,
                                                double& X0, double& X1, double& X2) {
    const double dd = (double)dep;
    const double q0 = dd * bx - pose[9], q1 = dd * by - pose[10], q2 = dd * bz - pose[11];
    X0 = pose[0] * q0 + pose[3] * q1 + pose[6] * q2;
    X1 = pose[1] * q0 + pose[4] * q1 + pose[7] * q2;
    X2 = pose[2] * q0 + pose[5] * q1 + pose[8] * q2;
}

// Backproject_equirect (depthmap.cc:106-111): R^T (d b - t), b from the bearing table
__device__ __forceinline__ void backproject(const GeoArgs& g, const double* pose, int i, int j, float dep,
                                            double& X0, double& X1, double& X2) {
    const float2 r = g.rowtab[i], c = g.coltab[j];
    backproject_dir(pose, (double)(r.x * c.x), (double)(-r.y), (double)(r.x * c.y), dep, X0, X1, X2);
}

// Bearing at a continuous pixel position (x, y), x in [0, W), y in [0, H):
// (cos phi sin lambda, -sin phi, cos phi cos lambda) with lambda = 2 pi (x/W - 1/2)
// and phi = pi (1/2 - y/H), the bearing table's formula; iw = 1/W, ih = 1/H.
// The angles go to sincos2pi as turns: x/W + 1/2 (mod 1) and 1/4 - y/(2H) (mod 1).
__device__ __forceinline__ void bearing_at(float x, float y, float iw, float ih, float& bx, float& by, float& bz) {
    float ul = fma_(x, iw, 0.5f);
    ul = ul - floor_(ul);
    float up = fma_(y, -0.5f * ih, 0.25f);
    if (up < 0.0f) up = up + 1.0f;
    float sl, cl, sp, cp;
    sincos2pi(ul, &sl, &cl);
    sincos2pi(up, &sp, &cp);
    bx = cp * sl;
    by = -sp;
    bz = cp * cl;
}

// ReprojectEquirect (depthmap.cc:143-161) plus the caller's distance
// cv::norm(R X + t); returns true when too near (|R X + t| < 1e-4).
__device__ __forceinline__ bool reproject(const GeoArgs& g, const double* pose, double X0, double X1, double X2,
                                          float& x, float& y, float& dist) {
    const double c0 = pose[0] * X0 + pose[1] * X1 + pose[2] * X2 + pose[9];
    const double c1 = pose[3] * X0 + pose[4] * X1 + pose[5] * X2 + pose[10];
    const double c2 = pose[6] * X0 + pose[7] * X1 + pose[8] * X2 + pose[11];
    const double r2 = c0 * c0 + c1 * c1 + c2 * c2;
    if (r2 < 1e-8) return true;
    const float xf = (float)c0, yf = (float)c1, zf = (float)c2;
    const float r2f = clampf(fma_(xf, xf, fma_(yf, yf, zf * zf)), 1e-30f, 1e30f);
    const float ir = rsqrt_pos(r2f);
    const float sy = clampf(yf * ir, -1.0f, 1.0f);
    y = fma_(asin_(sy), g.kv, g.hv);
    x = fma_(atan2_(xf, zf), g.ku, g.hu);
    dist = r2f * ir;
    return false;
}

// LinearInterpolation<float> (depthmap.cc:34-50): 0 outside [0, W-1) x [0, H-1)
__device__ __forceinline__ float lin_interp(const float* img, int W, int H, float y, float x) {
    if (x < 0.0f || x >= (float)(W - 1) || y < 0.0f || y >= (float)(H - 1)) return 0.0f;
    const int ix = (int)x, iy = (int)y;
    const float dx = x - (float)ix, dy = y - (float)iy;
    const float* r0 = img + (size_t)iy * W + ix;
    const float* r1 = r0 + W;
    const float im0 = (1.0f - dx) * r0[0] + dx * r0[1];
    const float im1 = (1.0f - dx) * r1[0] + dx * r1[1];
    return (1.0f - dy) * im0 + dy * im1;
}

}  // namespace pmp
