// pmp_math.h -- device elementary functions and random streams.
//
// Built only from IEEE-754 operations that are correctly rounded or exact on
// gfx950 (v_add/v_mul/v_fma_f32, compares, min/max/med3, v_floor_f32,
// v_fract_f32 of non-negative values, copysign, integer and bit ops): no
// division, no hardware sqrt/rcp/rsq approximations, no transcendental
// instructions.  With -ffp-contract=off the compiler emits exactly the fma /
// mul / add sequence written here, so results are reproducible bit for bit
// by any IEEE host executing the same sequence.
//
// Methods (published; coefficients fitted by gen_coeffs.py in this folder):
//   rcp    seed 0x7EF311C3 - bits(x), three Newton steps (<= 0.501 ulp)
//   rsqrt  seed 0x5F375A86 - (bits(x) >> 1) (Lomont 2003), one
//          fourth-order Householder step, one Newton step (<= 1 ulp)
//   asin   |x| <= 1/2: x + x^3 P(x^2); else pi/2 - 2 asin(sqrt((1-|x|)/2))
//          (Cody & Waite, Software Manual for the Elementary Functions, 1980)
//   acos   from the same core polynomial, three ranges
//   atan2  q = min/max, atan(q) = q P(q^2) on [0, 1], octant fix-up
//   lonlat r = rsqrt(h^2), h = sqrt(x^2 + z^2); longitude atan2(x, z) from
//          the octant angle asin(min(|x|, |z|) r) = s + s^3 P(s^2) on
//          [0, 1/sqrt 2]; latitude atan2(y, h) = atan(|y| r), pi/2 minus
//          atan(1/(|y| r)) above 45 degrees
//   exp    Cody-Waite reduction by ln2, degree-6 polynomial, 2^k by exponent bits
//   log    mantissa in [sqrt(1/2), sqrt 2), log1p(f) = f P(f)
//   sincos2pi  quadrant from floor(4u), sin/cos(pi f / 2) in f^2
//   Philox4x32-10  Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
//          easy as 1, 2, 3", SC'11
//   Box-Muller  Box & Muller, Ann. Math. Stat. 29 (1958)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PMP_DF __device__ __forceinline__

namespace pmp {

PMP_DF uint32_t f2u(float x) { return __float_as_uint(x); }
PMP_DF float u2f(uint32_t u) { return __uint_as_float(u); }
PMP_DF float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
PMP_DF float abs_(float x) { return __builtin_fabsf(x); }
PMP_DF float floor_(float x) { return __builtin_floorf(x); }

// clamp with NaN -> lo: v_max_f32 returns the non-NaN operand (IEEE mode), so
// min(max(x, lo), hi) equals (x >= lo) ? ((x <= hi) ? x : hi) : lo for every x
// (the two forms can differ only in the sign of a zero result when lo or hi
// is zero, which no caller distinguishes)
PMP_DF float clampf(float x, float lo, float hi) { return __builtin_fminf(__builtin_fmaxf(x, lo), hi); }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kPiO2 = 1.57079632679489661923f;

// 1/a, a in [2^-125, 2^125]
PMP_DF float rcp_pos(float a) {
    float y = u2f(0x7EF311C3u - f2u(a));
    float e;
    e = fma_(-a, y, 1.0f); y = fma_(y, e, y);
    e = fma_(-a, y, 1.0f); y = fma_(y, e, y);
    e = fma_(-a, y, 1.0f); y = fma_(y, e, y);
    return y;
}

// 1/x, |x| clamped to [1e-30, 1e30] (NaN -> 1e-30), sign kept
PMP_DF float rcp(float x) {
    float y = rcp_pos(clampf(abs_(x), 1e-30f, 1e30f));
    return (x < 0.0f) ? -y : y;
}

// 1/sqrt(a), a in [1e-30, 1e30]: one fourth-order step
// y (1 + r/2 + 3r^2/8 + 5r^3/16), r = 1 - a y^2, then one Newton step
PMP_DF float rsqrt_pos(float a) {
    float y = u2f(0x5F375A86u - (f2u(a) >> 1));
    float r = fma_(-(a * y), y, 1.0f);
    y = fma_(y * r, fma_(fma_(r, 0.3125f, 0.375f), r, 0.5f), y);
    const float ay = a * y;
    r = fma_(-ay, y, 1.0f);
    return fma_(0.5f * y, r, y);
}

// sqrt(x): x < 1e-30 (and NaN) -> 0, x > 1e30 -> 1e15
PMP_DF float sqrt_(float x) {
    const bool tiny = !(x >= 1e-30f);
    const float a = tiny ? 1.0f : ((x <= 1e30f) ? x : 1e30f);
    const float r = a * rsqrt_pos(a);
    return tiny ? 0.0f : r;
}

PMP_DF float asin_core(float a) {
    float z = a * a;
    float p = 4.221854731e-02f;
    p = fma_(p, z, 2.414761111e-02f);
    p = fma_(p, z, 4.547709599e-02f);
    p = fma_(p, z, 7.495241612e-02f);
    p = fma_(p, z, 1.666675359e-01f);
    return fma_(a * z, p, a);
}

// asin(x), x in [-1, 1]; one core polynomial, the square root only for |x| > 1/2
PMP_DF float asin_(float x) {
    const float a = abs_(x);
    const bool big = a > 0.5f;
    float in = a;
    if (big) in = sqrt_((1.0f - a) * 0.5f);
    const float c = asin_core(in);
    const float r = big ? fma_(-2.0f, c, kPiO2) : c;
    return (x < 0.0f) ? -r : r;
}

// acos(x), x in [-1, 1]
PMP_DF float acos_(float x) {
    if (x > 0.5f) {
        float s = sqrt_((1.0f - x) * 0.5f);
        return 2.0f * asin_core(s);
    }
    if (x < -0.5f) {
        float s = sqrt_((1.0f + x) * 0.5f);
        return fma_(-2.0f, asin_core(s), kPi);
    }
    return kPiO2 - asin_core(abs_(x)) * ((x < 0.0f) ? -1.0f : 1.0f);
}

PMP_DF float atan_core(float q) {
    float z = q * q;
    float p = 2.903517568e-03f;
    p = fma_(p, z, -1.628286950e-02f);
    p = fma_(p, z, 4.303914309e-02f);
    p = fma_(p, z, -7.533657551e-02f);
    p = fma_(p, z, 1.065466851e-01f);
    p = fma_(p, z, -1.420713216e-01f);
    p = fma_(p, z, 1.999305338e-01f);
    p = fma_(p, z, -3.333309293e-01f);
    p = fma_(p, z, 1.0f);
    return q * p;
}

// min/max of two non-negative floats (no signed zeros to tell apart); a NaN
// operand gives the other one (v_min_f32 / v_max_f32, C fminf / fmaxf)
PMP_DF float min_(float a, float b) { return __builtin_fminf(a, b); }
PMP_DF float max_(float a, float b) { return __builtin_fmaxf(a, b); }
// median of three (v_med3_f32; with a NaN operand the minimum of the others).
// med3(a, b, 0) is min(a, b) and med3(a, b, K) is max(a, b) for a, b in
// [0, K]; unlike fminf/fmaxf it needs no quieting of its operands.
PMP_DF float med3_(float a, float b, float c) { return __builtin_amdgcn_fmed3f(a, b, c); }

// atan2(y, x) for |y|, |x| <= 1e30, in [-pi, pi], with the sign of y
// (atan2(+-0, x >= 0) = +-0, atan2(+-0, x < 0) = +-pi).  The divisor max(|y|,
// |x|) is floored at 1e-30 and, within that range, needs no upper clamp for
// rcp_pos (valid up to 2^125).  The ratio min/max
// needs no clamp to [0, 1]: rcp_pos's seed and steps scale exactly with the
// binade, m rcp_pos(m) <= 1 for every float m in [1, 2) (checked exhaustively),
// and rounding is monotone, so mn rcp_pos(mx) <= mx rcp_pos(mx) <= 1.
PMP_DF float atan2_(float y, float x) {
    const float ay = abs_(y), ax = abs_(x);
    const float mn = med3_(ay, ax, 0.0f);                           // min of two non-negatives
    const float mx = max_(max_(ay, ax), 1e-30f);                    // max, floored (one v_max3_f32)
    const float q = mn * rcp_pos(mx);
    float a = atan_core(q);
    if (ay > ax) a = kPiO2 - a;
    if (x < 0.0f) a = kPi - a;
    return __builtin_copysignf(a, y);
}

// asin(s) for s in [0, 1/sqrt 2] (the octant angle of the longitude):
// s + s^3 P(s^2)
PMP_DF float asin_oct(float s) {
    const float z = s * s;
    float p = 1.016279534e-01f;
    p = fma_(p, z, -5.578331649e-02f);
    p = fma_(p, z, 6.111577526e-02f);
    p = fma_(p, z, 3.886268660e-02f);
    p = fma_(p, z, 7.548701018e-02f);
    p = fma_(p, z, 1.666531712e-01f);
    return fma_(s * z, p, s);
}

// Longitude atan2(x, z) and latitude atan2(y, h), h = sqrt(x^2 + z^2), of a
// direction with finite components (|x|, |y|, |z| < 1e7), from one
// r = rsqrt(h^2) (h^2 clamped to [1e-30, 1e30]):
//   longitude: the angle from the nearer of the z and x axes is
//     asin(min(|x|, |z|) r), in [0, pi/4]; octant fix-up as atan2_;
//   latitude: atan(t), t = |y| r, or pi/2 - atan(1/t) above 45 degrees
//     (t > 1).  The ratio in [0, 1] is min(t, rcp(t)) with no branch: the
//     reciprocal is below 1 for t > 1 and at least 1 for t < 1, and for t = 0
//     (rcp_pos gives NaN) the minimum is t (v_min_f32 returns the non-NaN
//     operand).
// x = z = 0 gives the longitude +-0 and the latitude +-pi/2 (t = |y| 1e15);
// y = 0 the latitude +-0.  Below h^2 = 1e-30 (only within 1e-15 of a pole)
// the clamped r makes the longitude approximate.
PMP_DF void lonlat(float x, float y, float z, float& lon, float& lat) {
    const float ax = abs_(x), az = abs_(z);
    const float h2 = fma_(x, x, z * z);
    const float r = rsqrt_pos(med3_(h2, 1e-30f, 1e30f));
    float a = asin_oct(med3_(ax, az, 0.0f) * r);
    if (ax > az) a = kPiO2 - a;
    if (z < 0.0f) a = kPi - a;
    lon = __builtin_copysignf(a, x);
    const float t = abs_(y) * r;
    float b = atan_core(min_(t, rcp_pos(t)));
    if (t > 1.0f) b = kPiO2 - b;
    lat = __builtin_copysignf(b, y);
}

// exp(x); x <= -87.3 or NaN -> 0; x >= 88 clamped
PMP_DF float exp_(float x) {
    if (!(x > -87.3f)) return 0.0f;
    x = (x < 88.0f) ? x : 88.0f;
    float kf = floor_(fma_(x, 1.44269504088896341f, 0.5f));
    float r = fma_(-kf, 6.93145752e-1f, x);
    r = fma_(-kf, 1.42860677e-6f, r);
    float p = 1.395605155e-03f;
    p = fma_(p, r, 8.375128731e-03f);
    p = fma_(p, r, 4.166608304e-02f);
    p = fma_(p, r, 1.666641533e-01f);
    p = fma_(p, r, 5.000000000e-01f);
    p = fma_(p, r, 1.0f);
    p = fma_(p, r, 1.0f);
    int k = (int)kf;
    int k1 = k >> 1, k2 = k - k1;
    p = p * u2f((uint32_t)(k1 + 127) << 23);
    return p * u2f((uint32_t)(k2 + 127) << 23);
}

// log(x), x in (0, 1] of the form (n+1) 2^-24
PMP_DF float log_(float x) {
    uint32_t b = f2u(x);
    int e = (int)(b >> 23) - 127;
    float m = u2f((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    float f = m - 1.0f;
    float p = -7.830716670e-02f;
    p = fma_(p, f, 1.301909983e-01f);
    p = fma_(p, f, -1.321227849e-01f);
    p = fma_(p, f, 1.416073143e-01f);
    p = fma_(p, f, -1.661034077e-01f);
    p = fma_(p, f, 2.000320554e-01f);
    p = fma_(p, f, -2.500157058e-01f);
    p = fma_(p, f, 3.333330154e-01f);
    p = fma_(p, f, -4.999998808e-01f);
    p = fma_(p, f, 1.0f);
    return fma_((float)e, 6.93147182e-1f, f * p);
}

// sin and cos of 2 pi u, u in [0, 1)
PMP_DF void sincos2pi(float u, float* s_out, float* c_out) {
    float w = u * 4.0f;
    float qf = floor_(w);
    float f = w - qf;
    int q = (int)qf & 3;
    float z = f * f;
    float sp = -3.431869573e-06f;
    sp = fma_(sp, z, 1.602547854e-04f);
    sp = fma_(sp, z, -4.681658000e-03f);
    sp = fma_(sp, z, 7.969260216e-02f);
    sp = fma_(sp, z, -6.459640861e-01f);
    sp = fma_(sp, z, 1.570796371e+00f);
    float s = f * sp;
    float cp = -2.382491766e-05f;
    cp = fma_(cp, z, 9.177248576e-04f);
    cp = fma_(cp, z, -2.086268924e-02f);
    cp = fma_(cp, z, 2.536693215e-01f);
    cp = fma_(cp, z, -1.233700514e+00f);
    cp = fma_(cp, z, 1.0f);
    float c = cp;
    float so = (q == 0) ? s : (q == 1) ? c : (q == 2) ? -s : -c;
    float co = (q == 0) ? c : (q == 1) ? -s : (q == 2) ? -c : s;
    *s_out = so;
    *c_out = co;
}

// ---- Philox4x32-10 ----
struct u32x4 { uint32_t x, y, z, w; };

PMP_DF u32x4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {c0, c1, c2, c3};
}

PMP_DF float u01_oc(uint32_t x) { return (float)((x >> 8) + 1u) * 5.9604644775390625e-8f; }
PMP_DF float u01_co(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-8f; }

PMP_DF void box_muller(uint32_t a, uint32_t b, float* z0, float* z1) {
    float u1 = u01_oc(a), u2 = u01_co(b);
    float r = sqrt_(-2.0f * log_(u1));
    float s, c;
    sincos2pi(u2, &s, &c);
    *z0 = r * c;
    *z1 = r * s;
}

constexpr uint32_t kStreamInit = 0x494E4954u;    // "INIT"
constexpr uint32_t kStreamRefine = 0x52464E45u;  // "RFNE"

}  // namespace pmp
