/*
 * or_math.h -- elementary functions and random streams of the CPU oracle.
 *
 * TEST INFRASTRUCTURE.  This file belongs to the oracle: the CPU restatement
 * that checks the HIP path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never includes it.
 *
 * Every function here is built from IEEE-754 operations that are correctly
 * rounded on any conforming machine (add, sub, mul, fma, compare/select,
 * floor, integer and bit operations).  No libm call, no division, no sqrt.
 * Compiled with -ffp-contract=off, the same sequence of operations gives the
 * same bits on an x86 host and on a gfx950 GPU; the device copy of these
 * functions (patchmatch_amd/csrc/pmp_math.h) performs the identical sequence,
 * and the GPU test suite checks the two bit for bit.
 *
 * Methods (published; coefficients are this project's own least-squares /
 * Lawson fits, produced by patchmatch_amd/csrc/gen_coeffs.py):
 *   rcp   : seed 0x7EF311C3 - bits(x), three Newton steps e=1-x*y, y+=y*e
 *           (max error 0.501 ulp on [1,2), exhaustive check).
 *   rsqrt : seed 0x5F375A86 - (bits(x)>>1) (C. Lomont, "Fast inverse square
 *           root", 2003), one fourth-order Householder step and one Newton
 *           step in residual form (<= 1 ulp, exhaustive over [1, 4)).
 *   asin  : |x|<=1/2: x + x^3 P(x^2); else pi/2 - 2 asin(sqrt((1-|x|)/2))
 *           (the classical reduction, e.g. Cody & Waite 1980, ch. 10).
 *   acos  : pi/2 - asin(x) for |x|<=1/2, 2 asin(sqrt((1-x)/2)) for x>1/2,
 *           pi - 2 asin(sqrt((1+x)/2)) for x<-1/2.
 *   atan2 : q = min/max in [0,1], atan(q) = q P(q^2), octant fix-up,
 *           sign of y.
 *   lonlat: r = rsqrt(h^2), h = sqrt(x^2+z^2); longitude atan2(x, z) from
 *           the octant angle asin(min(|x|,|z|) r) = s + s^3 P(s^2) on
 *           [0, 1/sqrt 2]; latitude atan2(y, h) = atan(|y| r), pi/2 minus
 *           atan(1/(|y| r)) above 45 degrees, through atan's polynomial.
 *   exp   : Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, degree-6
 *           polynomial, scale by 2^k through the exponent field.
 *   log   : x = m 2^e, m in [sqrt(1/2), sqrt 2), log1p(m-1) = f P(f).
 *   sincos2pi(u): angle 2 pi u; quadrant from floor(4u), sin/cos(pi/2 f) by
 *           polynomials in f^2 on [0,1].
 *   Philox4x32-10: J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
 *           "Parallel random numbers: as easy as 1, 2, 3", SC'11.
 *   Box-Muller: G. E. P. Box, M. E. Muller, Ann. Math. Stat. 29 (1958).
 */
#ifndef OR_MATH_H
#define OR_MATH_H

#include <stdint.h>
#include <string.h>
#include <math.h> /* fmaf, floorf, fabsf, copysignf only: all exact or correctly rounded */

static inline uint32_t or_f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float or_u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

/* clamp with NaN -> lo; comparisons only, no fminf/fmaxf (their signed-zero
 * behaviour is unspecified) */
static inline float or_clamp(float x, float lo, float hi) {
    return (x >= lo) ? ((x <= hi) ? x : hi) : lo;
}

#define OR_PI_F      3.14159265358979323846f
#define OR_PIO2_F    1.57079632679489661923f

/* 1/a for a in [2^-125, 2^125] (callers clamp) */
static inline float or_rcp_pos(float a) {
    float y = or_u2f(0x7EF311C3u - or_f2u(a));
    float e;
    e = fmaf(-a, y, 1.0f); y = fmaf(y, e, y);
    e = fmaf(-a, y, 1.0f); y = fmaf(y, e, y);
    e = fmaf(-a, y, 1.0f); y = fmaf(y, e, y);
    return y;
}

/* 1/x for any x: |x| clamped to [1e-30, 1e30] (NaN -> 1e-30), sign kept
 * (x < 0 gives a negative result; -0 and NaN give positive) */
static inline float or_rcp(float x) {
    float a = or_clamp(fabsf(x), 1e-30f, 1e30f);
    float y = or_rcp_pos(a);
    return (x < 0.0f) ? -y : y;
}

/* 1/sqrt(a) for a in [1e-30, 1e30] (callers clamp): one fourth-order step
 * y (1 + r/2 + 3r^2/8 + 5r^3/16), r = 1 - a y^2, then one Newton step */
static inline float or_rsqrt_pos(float a) {
    float y = or_u2f(0x5F375A86u - (or_f2u(a) >> 1));
    float r = fmaf(-(a * y), y, 1.0f);
    y = fmaf(y * r, fmaf(fmaf(r, 0.3125f, 0.375f), r, 0.5f), y);
    float ay = a * y;
    r = fmaf(-ay, y, 1.0f);
    return fmaf(0.5f * y, r, y);
}

/* sqrt(x) for x >= 0: x < 1e-30 (and NaN) -> 0, x > 1e30 -> 1e15 */
static inline float or_sqrt(float x) {
    if (!(x >= 1e-30f)) return 0.0f;
    float a = (x <= 1e30f) ? x : 1e30f;
    return a * or_rsqrt_pos(a);
}

/* asin(a) + correction polynomial, a in [0, 1/2] */
static inline float or_asin_core(float a) {
    float z = a * a;
    float p = 4.221854731e-02f;
    p = fmaf(p, z, 2.414761111e-02f);
    p = fmaf(p, z, 4.547709599e-02f);
    p = fmaf(p, z, 7.495241612e-02f);
    p = fmaf(p, z, 1.666675359e-01f);
    return fmaf(a * z, p, a);
}

/* asin(x) for x in [-1, 1] (callers clamp) */
static inline float or_asin(float x) {
    float a = fabsf(x);
    float r;
    if (a <= 0.5f) {
        r = or_asin_core(a);
    } else {
        float z = (1.0f - a) * 0.5f;           /* exact (Sterbenz, power of 2) */
        float s = or_sqrt(z);
        r = fmaf(-2.0f, or_asin_core(s), OR_PIO2_F);
    }
    return (x < 0.0f) ? -r : r;
}

/* acos(x) for x in [-1, 1] (callers clamp) */
static inline float or_acos(float x) {
    if (x > 0.5f) {
        float s = or_sqrt((1.0f - x) * 0.5f);
        return 2.0f * or_asin_core(s);
    }
    if (x < -0.5f) {
        float s = or_sqrt((1.0f + x) * 0.5f);
        return fmaf(-2.0f, or_asin_core(s), OR_PI_F);
    }
    return OR_PIO2_F - or_asin_core(fabsf(x)) * ((x < 0.0f) ? -1.0f : 1.0f);
}

/* atan(q) for q in [0, 1] */
static inline float or_atan_core(float q) {
    float z = q * q;
    float p = 2.903517568e-03f;
    p = fmaf(p, z, -1.628286950e-02f);
    p = fmaf(p, z, 4.303914309e-02f);
    p = fmaf(p, z, -7.533657551e-02f);
    p = fmaf(p, z, 1.065466851e-01f);
    p = fmaf(p, z, -1.420713216e-01f);
    p = fmaf(p, z, 1.999305338e-01f);
    p = fmaf(p, z, -3.333309293e-01f);
    p = fmaf(p, z, 1.0f);
    return q * p;
}

/* min/max of two non-negative floats (no signed zeros to tell apart) with
 * IEEE minNum/maxNum NaN handling: a NaN operand gives the other one, as
 * v_min_f32 / v_max_f32 on the GPU */
static inline float or_min(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    return (b < a) ? b : a;
}
static inline float or_max(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    return (b > a) ? b : a;
}

/* median of three (v_med3_f32): with a NaN operand, the minimum of the
 * others (IEEE-mode min ignores a quiet NaN); no caller passes signed zeros
 * that the median would have to order */
static inline float or_med3(float a, float b, float c) {
    if (a != a || b != b || c != c) return or_min(or_min(a, b), c);
    float lo = (b < a) ? b : a, hi = (b < a) ? a : b;
    float m = (c < hi) ? c : hi;
    return (lo < m) ? m : lo;
}

/* atan2(y, x) for |y|, |x| <= 1e30; result in [-pi, pi] with the sign of y
 * (atan2(+-0, x >= 0) = +-0, atan2(+-0, x < 0) = +-pi).  mn / mx is not
 * clamped: x or_rcp_pos(x) <= 1 for every x in [1, 2), the seed and steps scale
 * exactly with the binade, and rounding is monotone (pinned by the tests). */
static inline float or_atan2(float y, float x) {
    float ay = fabsf(y), ax = fabsf(x);
    float mn = or_med3(ay, ax, 0.0f);                              /* min of two non-negatives */
    float mx = or_max(or_max(ay, ax), 1e-30f);                     /* max, floored */
    float q = mn * or_rcp_pos(mx);                                 /* in [0, 1] */
    float a = or_atan_core(q);
    if (ay > ax) a = OR_PIO2_F - a;
    if (x < 0.0f) a = OR_PI_F - a;
    return copysignf(a, y);
}

/* asin(s) for s in [0, 1/sqrt 2] (the octant angle of the longitude):
 * s + s^3 P(s^2) */
static inline float or_asin_oct(float s) {
    float z = s * s;
    float p = 1.016279534e-01f;
    p = fmaf(p, z, -5.578331649e-02f);
    p = fmaf(p, z, 6.111577526e-02f);
    p = fmaf(p, z, 3.886268660e-02f);
    p = fmaf(p, z, 7.548701018e-02f);
    p = fmaf(p, z, 1.666531712e-01f);
    return fmaf(s * z, p, s);
}

/* longitude atan2(x, z) and latitude atan2(y, h), h = sqrt(x^2 + z^2), of a
 * direction with finite components (|x|, |y|, |z| < 1e7), from one
 * r = rsqrt(h^2) (h^2 clamped to [1e-30, 1e30]):
 *   longitude: the angle from the nearer of the z and x axes is
 *     asin(min(|x|, |z|) r), in [0, pi/4]; octant fix-up as atan2;
 *   latitude: atan(t), t = |y| r, or pi/2 - atan(1/t) above 45 degrees
 *     (t > 1).  The ratio in [0, 1] is min(t, rcp(t)): the reciprocal is
 *     below 1 for t > 1 and at least 1 for t < 1, and for t = 0 (or_rcp_pos
 *     gives NaN) or_min returns t.
 * x = z = 0 gives the longitude +-0 and the latitude +-pi/2 (t = |y| 1e15);
 * y = 0 the latitude +-0.  Below h^2 = 1e-30 (only within 1e-15 of a pole)
 * the clamped r makes the longitude approximate. */
static inline void or_lonlat(float x, float y, float z, float* lon, float* lat) {
    float ax = fabsf(x), az = fabsf(z);
    float h2 = fmaf(x, x, z * z);
    float r = or_rsqrt_pos(or_med3(h2, 1e-30f, 1e30f));
    float a = or_asin_oct(or_med3(ax, az, 0.0f) * r);
    if (ax > az) a = OR_PIO2_F - a;
    if (z < 0.0f) a = OR_PI_F - a;
    *lon = copysignf(a, x);
    float t = fabsf(y) * r;
    float b = or_atan_core(or_min(t, or_rcp_pos(t)));
    if (t > 1.0f) b = OR_PIO2_F - b;
    *lat = copysignf(b, y);
}

/* exp(x); x <= -87.3 -> 0, x >= 88 -> e^88 clamp; NaN -> e^-87.3 path (0) */
static inline float or_exp(float x) {
    if (!(x > -87.3f)) return 0.0f;
    x = (x < 88.0f) ? x : 88.0f;
    float kf = floorf(fmaf(x, 1.44269504088896341f, 0.5f));
    float r = fmaf(-kf, 6.93145752e-1f, x);          /* ln2 high part (exact product) */
    r = fmaf(-kf, 1.42860677e-6f, r);                /* ln2 low part */
    float p = 1.395605155e-03f;
    p = fmaf(p, r, 8.375128731e-03f);
    p = fmaf(p, r, 4.166608304e-02f);
    p = fmaf(p, r, 1.666641533e-01f);
    p = fmaf(p, r, 5.000000000e-01f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int k = (int)kf;                                  /* in [-126, 127] */
    int k1 = k >> 1, k2 = k - k1;                     /* 2^k = 2^k1 * 2^k2, both normal */
    p = p * or_u2f((uint32_t)(k1 + 127) << 23);
    return p * or_u2f((uint32_t)(k2 + 127) << 23);
}

/* natural log for x in (0, 1] given as u = (n + 1) 2^-24, n < 2^24 (normal) */
static inline float or_log(float x) {
    uint32_t b = or_f2u(x);
    int e = (int)(b >> 23) - 127;
    float m = or_u2f((b & 0x007FFFFFu) | 0x3F800000u);   /* [1, 2) */
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    float f = m - 1.0f;                                   /* exact */
    float p = -7.830716670e-02f;
    p = fmaf(p, f, 1.301909983e-01f);
    p = fmaf(p, f, -1.321227849e-01f);
    p = fmaf(p, f, 1.416073143e-01f);
    p = fmaf(p, f, -1.661034077e-01f);
    p = fmaf(p, f, 2.000320554e-01f);
    p = fmaf(p, f, -2.500157058e-01f);
    p = fmaf(p, f, 3.333330154e-01f);
    p = fmaf(p, f, -4.999998808e-01f);
    p = fmaf(p, f, 1.0f);
    return fmaf((float)e, 6.93147182e-1f, f * p);
}

/* sin and cos of 2 pi u for u in [0, 1) */
static inline void or_sincos2pi(float u, float* s_out, float* c_out) {
    float w = u * 4.0f;                      /* exact */
    float qf = floorf(w);
    float f = w - qf;                        /* exact, [0,1) */
    int q = (int)qf & 3;
    float z = f * f;
    float sp = -3.431869573e-06f;
    sp = fmaf(sp, z, 1.602547854e-04f);
    sp = fmaf(sp, z, -4.681658000e-03f);
    sp = fmaf(sp, z, 7.969260216e-02f);
    sp = fmaf(sp, z, -6.459640861e-01f);
    sp = fmaf(sp, z, 1.570796371e+00f);
    float s = f * sp;
    float cp = -2.382491766e-05f;
    cp = fmaf(cp, z, 9.177248576e-04f);
    cp = fmaf(cp, z, -2.086268924e-02f);
    cp = fmaf(cp, z, 2.536693215e-01f);
    cp = fmaf(cp, z, -1.233700514e+00f);
    cp = fmaf(cp, z, 1.0f);
    float c = cp;
    float so, co;
    switch (q) {
        case 0: so = s;  co = c;  break;
        case 1: so = c;  co = -s; break;
        case 2: so = -s; co = -c; break;
        default: so = -c; co = s; break;
    }
    *s_out = so; *c_out = co;
}

/* ---------------- Philox4x32-10 ---------------- */
static inline void or_philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* uniform in (0, 1]: ((x >> 8) + 1) 2^-24 */
static inline float or_u01_oc(uint32_t x) { return (float)((x >> 8) + 1u) * 5.9604644775390625e-8f; }
/* uniform in [0, 1): (x >> 8) 2^-24 */
static inline float or_u01_co(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-8f; }

/* Box-Muller pair from two words */
static inline void or_box_muller(uint32_t a, uint32_t b, float* z0, float* z1) {
    float u1 = or_u01_oc(a), u2 = or_u01_co(b);
    float r = or_sqrt(-2.0f * or_log(u1));
    float s, c;
    or_sincos2pi(u2, &s, &c);
    *z0 = r * c;
    *z1 = r * s;
}

/* stream tags of the counter's last word */
#define OR_STREAM_INIT   0x494E4954u   /* "INIT" */
#define OR_STREAM_REFINE 0x52464E45u   /* "RFNE" */

#endif
