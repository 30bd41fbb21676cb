// synth_render.hip -- GPU ray caster of the seeded synthetic scenes
// (patchmatch_amd/synth.py), for benchmark inputs at sizes the numpy renderer
// cannot produce in seconds.  Not part of the densifier (libpmp.so): it is a
// separate library, libpmpsynth.so.
//
// Same scene description and convention as synth.Scene / synth.render, in
// double precision: box room [-4,4]x[-1.5,1.5]x[-3,3] (faces give object ids
// 0..5), spheres (ids 6..), an open tube along x (the next id), 4-octave value
// noise with a quintic fade (128 + 110 * noise, clamped to [0, 255]).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace {

struct SceneArgs {
    double R[9], C[3];            // world->camera rotation (row-major), camera centre (world)
    const double* spheres;        // n x (cx, cy, cz, r)
    int n_spheres;
    int has_tube;
    double tube[5];               // y0, z0, radius, x0, x1
    const int32_t* perm;          // 2 * 256
    const double* vals;           // 256
    int mask;
    double base_freq;
    int W, H;
};

__device__ double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

__device__ double noise(const SceneArgs& s, double x, double y, double z) {
    const double fx0 = floor(x), fy0 = floor(y), fz0 = floor(z);
    const long long ix = (long long)fx0, iy = (long long)fy0, iz = (long long)fz0;
    const double ux = fade(x - fx0), uy = fade(y - fy0), uz = fade(z - fz0);
    const int m = s.mask;
    double out = 0.0;
    for (int dx = 0; dx < 2; ++dx) {
        const double wx = dx ? ux : 1.0 - ux;
        const long long hx = s.perm[(ix + dx) & m];
        for (int dy = 0; dy < 2; ++dy) {
            const double wy = dy ? uy : 1.0 - uy;
            const long long hy = s.perm[(hx + iy + dy) & m];
            for (int dz = 0; dz < 2; ++dz) {
                const double wz = dz ? uz : 1.0 - uz;
                out += wx * wy * wz * s.vals[s.perm[(hy + iz + dz) & m]];
            }
        }
    }
    return out;
}

__global__ void k_render(SceneArgs s, float* __restrict__ img, float* __restrict__ depth) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= s.W || i >= s.H) return;
    const double lon = 2.0 * M_PI * ((double)j / s.W - 0.5);
    const double lat = -M_PI * ((double)i / s.H - 0.5);
    const double b0 = cos(lat) * sin(lon), b1 = -sin(lat), b2 = cos(lat) * cos(lon);
    double D[3];
    for (int k = 0; k < 3; ++k) D[k] = b0 * s.R[0 * 3 + k] + b1 * s.R[1 * 3 + k] + b2 * s.R[2 * 3 + k];
    const double lo[3] = {-4.0, -1.5, -3.0}, hi[3] = {4.0, 1.5, 3.0};
    double t = INFINITY;
    int obj = 0;
    for (int k = 0; k < 3; ++k) {
        const double tb = D[k] > 0 ? (hi[k] - s.C[k]) / D[k] : (lo[k] - s.C[k]) / D[k];
        if (isfinite(tb) && tb > 0 && tb < t) {
            t = tb;
            obj = k * 2 + (D[k] > 0 ? 1 : 0);
        }
    }
    for (int q = 0; q < s.n_spheres; ++q) {
        const double* sp = s.spheres + 4 * q;
        const double o0 = s.C[0] - sp[0], o1 = s.C[1] - sp[1], o2 = s.C[2] - sp[2];
        const double b = D[0] * o0 + D[1] * o1 + D[2] * o2;
        const double c = o0 * o0 + o1 * o1 + o2 * o2 - sp[3] * sp[3];
        const double disc = b * b - c;
        if (disc > 0) {
            const double sq = sqrt(disc);
            const double t1 = -b - sq, t2 = -b + sq;
            const double ts = t1 > 1e-9 ? t1 : (t2 > 1e-9 ? t2 : INFINITY);
            if (ts < t) {
                t = ts;
                obj = 6 + q;
            }
        }
    }
    if (s.has_tube) {
        const double y0 = s.tube[0], z0 = s.tube[1], Rr = s.tube[2], x0 = s.tube[3], x1 = s.tube[4];
        const double oy = s.C[1] - y0, oz = s.C[2] - z0;
        const double a = D[1] * D[1] + D[2] * D[2];
        const double b = D[1] * oy + D[2] * oz;
        const double c = oy * oy + oz * oz - Rr * Rr;
        const double disc = b * b - a * c;
        if (disc > 0 && a > 1e-12) {
            const double sq = sqrt(disc);
            double best = INFINITY;
            const double roots[2] = {(-b - sq) / a, (-b + sq) / a};
            for (int r = 0; r < 2; ++r) {
                const double xh = s.C[0] + roots[r] * D[0];
                if (roots[r] > 1e-9 && xh >= x0 && xh <= x1 && roots[r] < best) best = roots[r];
            }
            if (best < t) {
                t = best;
                obj = 6 + s.n_spheres;
            }
        }
    }
    const double off = obj * 17.31;
    double P[3], val = 0.0, amp = 1.0, f = 1.0;
    for (int k = 0; k < 3; ++k) P[k] = (s.C[k] + t * D[k]) * s.base_freq + off;
    for (int o = 0; o < 4; ++o) {
        val += amp * noise(s, P[0] * f, P[1] * f, P[2] * f);
        amp *= 0.5;
        f *= 2.0;
    }
    double v = 128.0 + 110.0 * val;
    v = v < 0.0 ? 0.0 : (v > 255.0 ? 255.0 : v);
    const size_t k = (size_t)i * s.W + j;
    img[k] = (float)v;
    depth[k] = (float)t;
}

}  // namespace

extern "C" {

// Noise-free intensity and depth (host float arrays, H*W) of one camera.
// pose: world->camera R (row-major) and t; spheres: n x 4 doubles; tube: 5
// doubles or NULL; perm: 512 int32; vals: 256 doubles.  Returns 0 or a HIP
// error code.
int pmps_render(const double* pose, int W, int H, const double* spheres, int n_spheres, const double* tube,
                const int32_t* perm, const double* vals, int mask, double base_freq, float* img, float* depth) {
    SceneArgs s = {};
    for (int k = 0; k < 9; ++k) s.R[k] = pose[k];
    for (int r = 0; r < 3; ++r) {   // C = -R^T t
        double c = 0.0;
        for (int m = 0; m < 3; ++m) c -= pose[m * 3 + r] * pose[9 + m];
        s.C[r] = c;
    }
    s.n_spheres = n_spheres;
    s.has_tube = tube != nullptr;
    if (tube)
        for (int k = 0; k < 5; ++k) s.tube[k] = tube[k];
    s.mask = mask;
    s.base_freq = base_freq;
    s.W = W;
    s.H = H;
    const size_t P = (size_t)W * H;
    double* d_sph = nullptr;
    int32_t* d_perm = nullptr;
    double* d_vals = nullptr;
    float *d_img = nullptr, *d_dep = nullptr;
    hipError_t e = hipMalloc(&d_sph, (n_spheres > 0 ? n_spheres : 1) * 4 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&d_perm, 2 * (mask + 1) * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&d_vals, (mask + 1) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&d_img, P * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&d_dep, P * sizeof(float));
    if (e == hipSuccess && n_spheres > 0)
        e = hipMemcpy(d_sph, spheres, n_spheres * 4 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_perm, perm, 2 * (mask + 1) * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_vals, vals, (mask + 1) * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        s.spheres = d_sph;
        s.perm = d_perm;
        s.vals = d_vals;
        hipLaunchKernelGGL(k_render, dim3((W + 127) / 128, H), dim3(128), 0, 0, s, d_img, d_dep);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(img, d_img, P * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(depth, d_dep, P * sizeof(float), hipMemcpyDeviceToHost);
    void* bufs[] = {d_sph, d_perm, d_vals, d_img, d_dep};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    return (int)e;
}

}  // extern "C"
