// pmp_score.h -- device side of the plane-hypothesis score and its helpers.
//
// The score follows ComputeScorePlaneHypothesis (PatchMatchCuda.cu:46-180):
// for every tap (k1 outer, k2 inner) the reference bearing is intersected
// with the plane, moved into each neighbour camera, projected to
// equirectangular (u, v), sampled bilinearly, and accumulated into weighted
// moments; each view gives an NCC; views are aggregated.  The arithmetic is
// the canonical FP32 sequence of this project (identical, operation for
// operation, to the host restatement):
//   * bearings from separable row/column tables;
//   * per view, the direction sigma M b + c |b.p| of the plane point (no
//     per-tap reciprocal: the projection only needs the direction);
//   * longitude and latitude through one atan polynomial (lonlat);
//   * weights Wc[|dI|] * Ws[tap] from tables;
//   * shifted moments (reference taps centred on their weighted mean,
//     neighbour samples shifted by the view's centre-tap sample);
//   * PMP_AGG_REF_FOLD pools the moments across views as the reference does
//     (its sums are zeroed once, :57-62, outside the view loop :101), from
//     the per-view shifted moments (pool_view);
//   * column-pair packed uint8 neighbours (2 bytes per texel): one 32-bit
//     load per tap and view.
#pragma once
#include "pmp_internal.h"
#include "pmp_math.h"

namespace pmp {

// ---- coverage (PatchMatchCuda.cu:262,283; grids :465-470, :538-539, :578-581)
PMP_DF bool interior(const KfArgs& a, int i, int j) {
    int ph = a.half;
    return !(i < ph + 1 || i > a.H - ph - 1 || j < ph + 1 || j > a.W - ph - 1);
}
PMP_DF bool in_init_grid(const KfArgs& a, int i, int j) {
    if (a.coverage == PMP_COVER_FULL) return true;
    return i < (a.H / 16) * 16 && j < (a.W / 16) * 16;
}
PMP_DF bool in_score_grid(const KfArgs& a, int i, int j) {
    if (!interior(a, i, j)) return false;
    if (a.coverage == PMP_COVER_FULL) return true;
    return i < (a.H / 32) * 32 && j < (a.W / 32) * 32;
}
PMP_DF bool in_prop_grid(const KfArgs& a, int i, int j) {
    if (!interior(a, i, j)) return false;
    if (a.coverage == PMP_COVER_FULL) return true;
    int J = 32 * ((a.W / 32) / 2);
    return i < (a.H / 32) * 32 && (j + (j & 1)) < 2 * J;
}

PMP_DF void bearing(const KfArgs& a, int i, int j, float& bx, float& by, float& bz) {
    float2 r = a.rowtab[i];
    float2 c = a.coltab[j];
    bx = r.x * c.x;
    by = -r.y;
    bz = r.x * c.y;
}

struct RefStat {
    float Ic, mx, inv, mxc, vx;
};

// Plane-independent reference-side moments of the patch (two passes).
PMP_DF RefStat ref_stats(const KfArgs& a, const float* sWc, int i, int j) {
    RefStat s;
    const int W = a.W;
    s.Ic = (float)a.ref[(size_t)i * W + j];
    float Sw = 0.0f, Swx = 0.0f;
    int t = 0;
    for (int k1 = -a.half; k1 <= a.half; k1 += a.stride) {
        const uint8_t* row = a.ref + (size_t)(i + k1) * W + j;
        for (int k2 = -a.half; k2 <= a.half; k2 += a.stride, ++t) {
            float xs = (float)row[k2] - s.Ic;
            float w = sWc[(int)abs_(xs)] * a.Ws[t];
            Sw = Sw + w;
            Swx = fma_(w, xs, Swx);
        }
    }
    s.inv = rcp_pos(Sw);
    s.mx = Swx * s.inv;
    float Sc = 0.0f, Scc = 0.0f;
    t = 0;
    for (int k1 = -a.half; k1 <= a.half; k1 += a.stride) {
        const uint8_t* row = a.ref + (size_t)(i + k1) * W + j;
        for (int k2 = -a.half; k2 <= a.half; k2 += a.stride, ++t) {
            float xs = (float)row[k2] - s.Ic;
            float w = sWc[(int)abs_(xs)] * a.Ws[t];
            float xc = xs - s.mx;
            float wxc = w * xc;
            Sc = Sc + wxc;
            Scc = fma_(wxc, xc, Scc);
        }
    }
    s.mxc = Sc * s.inv;
    s.vx = fma_(-s.mxc, s.mxc, Scc * s.inv);
    return s;
}

PMP_DF float ncc(const RefStat& rs, float floor_, float Swy, float Swyy, float Swxy) {
    float my = Swy * rs.inv;
    float vy = fma_(-my, my, Swyy * rs.inv);
    float cov = fma_(-rs.mxc, my, Swxy * rs.inv);
    if (rs.vx < floor_ || vy < floor_) return -1.0f;
    float s = cov * rsqrt_pos(rs.vx * vy);
    return clampf(s, -1.0f, 1.0f);
}

// Neighbour sample at (u, v); texel k at integer k (PatchMatchCuda.cu:146 looks
// up (u+0.5, v+0.5) with texel centres at k+0.5).  A view is packed as rows
// of W + 2 halfwords, halfword (x, y) = p(x, y) | p(x, y+1) << 8 (k_pack); the
// 4 bytes at halfword (x, y), a 2-byte-aligned 32-bit load with a uniform
// 64-bit base and a 32-bit byte offset, are the sample's 2x2 neighbourhood.
// Two bytes per texel instead of four halve the views' cache footprint.  The
// clamps below keep every load inside the view (a NaN coordinate clamps to 0).
//   kZero: CPU semantics (depthmap.cc:36-38), 0 outside [0,W-1) x [0,H-1);
//   else:  u clamped to [0, umax] (umax = W with longitude wrap: column W of
//          the halfwords repeats column 0, column W+1 column 1; W-1 for
//          clamp), v clamped to [0, H-1].
//   The instruction choices are exact for the clamped coordinates (uu, vv >= 0,
//   finite): truncation equals floor, fract(x) = x - floor(x) is exact, med3
//   with lo <= hi is the clamp (a NaN coordinate gives lo, as clampf), and
//   the 24-bit product covers rows and row strides below 2^24.
// A sample in flight: the 4 packed bytes being loaded and the weights.
template <bool kZero>
struct Pend {
    uint32_t word;
    float fx, fy;
    bool inside;   // kZero: the coordinate lies in [0,W-1) x [0,H-1)
};

// Issues the load of the sample at (u, v) (the blend waits for it).
template <bool kZero>
PMP_DF void gather(const KfArgs& a, const uint32_t* __restrict__ base, float u, float v, Pend<kZero>& p) {
    const int W = a.W;
    float uu, vv;
    bool inside = true;
    if constexpr (kZero) {
        inside = (u >= 0.0f) && (u < (float)(W - 1)) && (v >= 0.0f) && (v < a.vmax);
        uu = inside ? u : 0.0f;
        vv = inside ? v : 0.0f;
    } else {
        uu = __builtin_amdgcn_fmed3f(u, 0.0f, a.umax);
        vv = __builtin_amdgcn_fmed3f(v, 0.0f, a.vmax);
    }
    p.fx = __builtin_amdgcn_fractf(uu);
    p.fy = __builtin_amdgcn_fractf(vv);
    const uint32_t x0 = (uint32_t)(int)uu;
    const uint32_t y0 = (uint32_t)(int)vv;
    const uint32_t idx = __umul24(y0, (uint32_t)(W + 2)) + x0;
    typedef uint32_t u32_a2 __attribute__((aligned(2)));
    p.word = *(const u32_a2*)((const char*)base + idx * 2u);
    p.inside = inside;
}

template <bool kZero>
PMP_DF float blend(const Pend<kZero>& p) {
    const uint32_t word = p.word;
    const float b0 = (float)(word & 0xFFu);           // p(x, y)
    const float b1 = (float)((word >> 8) & 0xFFu);    // p(x, y+1)
    const float b2 = (float)((word >> 16) & 0xFFu);   // p(x+1, y)
    const float b3 = (float)(word >> 24);             // p(x+1, y+1)
    const float top = fma_(p.fx, b2 - b0, b0);
    const float bot = fma_(p.fx, b3 - b1, b1);
    const float val = fma_(p.fy, bot - top, top);
    if constexpr (kZero) return p.inside ? val : 0.0f;
    return val;
}

template <bool kZero>
PMP_DF float sample(const KfArgs& a, const uint32_t* __restrict__ base, float u, float v) {
    Pend<kZero> p;
    gather<kZero>(a, base, u, v, p);
    return blend<kZero>(p);
}

// Equirectangular projection (PatchMatchCuda.cu:132-144): the latitude
// asin(y / |X|) there is the same angle as atan2(y, sqrt(x^2 + z^2)) here.
// hu, hv are a.hu, a.hv (callers may pass copies held in vector registers).
PMP_DF void project(const KfArgs& a, float hu, float hv, float x, float y, float z, float& u, float& v) {
    float lon, lat;
    lonlat(x, y, z, lon, lat);
    u = fma_(lon, a.ku, hu);
    v = fma_(lat, a.kv, hv);
}

// A uniform value copied into a vector register, so that an fma with two
// other uniform operands needs no move of its own in a loop.
PMP_DF float in_vgpr(float s) {
    float v;
    asm("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
    return v;
}

// The fold of PatchMatchCuda.cu:171 with the agreement test written without
// the division (identical at s2 = 0).
PMP_DF float fold(float s, float s2, float agree) {
    if (abs_(s - s2) < agree * abs_(s2)) return 0.5f * (s + s2);
    return (s2 > s) ? s2 : s;
}

// Moments pooled across views (PMP_AGG_REF_FOLD).  The reference zeroes its
// six sums once (PatchMatchCuda.cu:57-62) and the view loop (:101-157) keeps
// adding to them, so view v's NCC (:161-169) is that of the taps of views
// 0..v together: the reference side repeats (its mean and variance stay those
// of one view), the pooled covariance is the mean of the per-view
// covariances, and the pooled variance of y is the mean of the per-view
// variances plus the spread of the per-view means.  C and V sum covariances
// and variances, mu and q are Welford's running mean and sum of squared
// deviations of the view means, s the fold.  Zero sums and s = -1 start it.
struct Pool {
    float C, V, mu, q, s;
};

// Adds view v (shifted moments acc0..2 about its shift) and folds the pooled
// NCC of views 0..v into st.s.
PMP_DF void pool_view(const KfArgs& a, const RefStat& rs, Pool& st, float rn, float acc0, float acc1, float acc2,
                      float shift) {
    const float my = acc0 * rs.inv;
    const float vy = fma_(-my, my, acc1 * rs.inv);
    const float cov = fma_(-rs.mxc, my, acc2 * rs.inv);
    const float m = shift + my;
    st.C = st.C + cov;
    st.V = st.V + vy;
    const float dm = m - st.mu;
    st.mu = fma_(dm, rn, st.mu);
    st.q = fma_(dm, m - st.mu, st.q);
    const float cp = st.C * rn;
    const float vp = (st.V + st.q) * rn;
    float s2 = -1.0f;
    if (!(rs.vx < a.var_floor || vp < a.var_floor)) s2 = clampf(cp * rsqrt_pos(rs.vx * vp), -1.0f, 1.0f);
    st.s = fold(st.s, s2, a.fold_agree);
}

// Aggregation of N (run-time) independent per-view scores kept in LDS at
// sv[v * ld] (v < N): FOLD_INDEP folds them in view order, MAX takes the
// largest, TOPK_MEAN inserts each into a descending list of the topk largest
// and sums that list in order.  Per-view scores are finite (ncc clamps to
// [-1, 1]).
PMP_DF float aggregate_rt(const KfArgs& a, const float* sv, int ld) {
    const int N = a.N;
    if (a.agg_mode == PMP_AGG_TOPK_MEAN) {
        float top[kMaxViews];
#pragma unroll
        for (int p = 0; p < kMaxViews; ++p) top[p] = -3.0e38f;
        for (int v = 0; v < N; ++v) {
            float x = sv[v * ld];
#pragma unroll
            for (int p = 0; p < kMaxViews; ++p) {
                if (p < a.topk) {
                    const float hi = (x > top[p]) ? x : top[p];
                    const float lo = (x > top[p]) ? top[p] : x;
                    top[p] = hi;
                    x = lo;
                }
            }
        }
        float sum = top[0];
#pragma unroll
        for (int p = 1; p < kMaxViews; ++p)
            if (p < a.topk) sum = sum + top[p];
        return sum * a.inv_k;
    }
    float s = -1.0f;
    if (a.agg_mode == PMP_AGG_MAX) {
        for (int v = 0; v < N; ++v) {
            const float s2 = sv[v * ld];
            s = (s2 > s) ? s2 : s;
        }
        return s;
    }
    for (int v = 0; v < N; ++v) s = fold(s, sv[v * ld], a.fold_agree);
    return s;
}

// Ray-plane intersection of tap bearing b with plane p (PatchMatchCuda.cu:117-122,
// |b.p| clamped to >= 1e-6) moved into a view: the camera-frame point b / s,
// s = b.p, lands at M b / s + c, whose direction sigma M b + c |s| (sigma =
// sign bit of s) is all the equirectangular projection uses.  tap_dir folds
// sigma into b and returns max(|s|, 1e-6); view_sample forms the direction
// for view v, projects and samples it.
PMP_DF void tap_dir(float bx, float by, float bz, float px, float py, float pz, float& sbx, float& sby, float& sbz,
                    float& as) {
    const float s = fma_(pz, bz, fma_(py, by, px * bx));
    const uint32_t sg = f2u(s) & 0x80000000u;
    sbx = u2f(f2u(bx) ^ sg);
    sby = u2f(f2u(by) ^ sg);
    sbz = u2f(f2u(bz) ^ sg);
    as = max_(abs_(s), 1e-6f);
}

template <bool kZero>
PMP_DF void view_gather(const KfArgs& a, float hu, float hv, const uint32_t* base, int v, float sbx, float sby,
                        float sbz, float as, Pend<kZero>& p) {
    const float x = fma_(a.c[v][0], as, fma_(a.M[v][0], sbx, fma_(a.M[v][1], sby, a.M[v][2] * sbz)));
    const float y = fma_(a.c[v][1], as, fma_(a.M[v][3], sbx, fma_(a.M[v][4], sby, a.M[v][5] * sbz)));
    const float z = fma_(a.c[v][2], as, fma_(a.M[v][6], sbx, fma_(a.M[v][7], sby, a.M[v][8] * sbz)));
    float u, vv;
    project(a, hu, hv, x, y, z, u, vv);
    gather<kZero>(a, base, u, vv, p);
}

template <bool kZero>
PMP_DF float view_sample(const KfArgs& a, float hu, float hv, const uint32_t* base, int v, float sbx, float sby,
                         float sbz, float as) {
    Pend<kZero> p;
    view_gather<kZero>(a, hu, hv, base, v, sbx, sby, sbz, as, p);
    return blend<kZero>(p);
}

// Where the per-tap inputs of the score come from.  GlobalTaps reads the
// bearing tables and the reference image in global memory; LdsTaps reads
// their copies for one propagation tile in LDS, so that waiting for them never
// waits for the neighbour gathers issued before them (vector-memory loads
// complete in order; LDS reads are counted separately).  rowoff(ii) + jj is
// the reference pixel (ii, jj); row offsets advance by rowstep() per row.
struct GlobalTaps {
    const float2* rowtab;
    const char* coltab;
    const uint8_t* ref;
    int W;
    PMP_DF float2 row(int ii) const { return rowtab[ii]; }
    PMP_DF float2 col(int jj) const { return *(const float2*)(coltab + (uint32_t)jj * 8u); }
    PMP_DF uint32_t rowoff(int ii) const { return (uint32_t)(ii * W); }
    PMP_DF uint32_t rowstep() const { return (uint32_t)W; }
    PMP_DF float refpx(uint32_t roff, int jj) const { return (float)ref[roff + (uint32_t)jj]; }
};
PMP_DF GlobalTaps global_taps(const KfArgs& a) { return GlobalTaps{a.rowtab, (const char*)a.coltab, a.ref, a.W}; }

// Tile copies: the tile's rows and columns plus a halo of half on each side,
// of the reference image (row stride kTileCols bytes) and of the bearing
// tables, r0 and c0 being the tile's first row and column minus half.
// kTileCols: 2 kTileSame tile columns, the largest halo each side, padded to a
// multiple of 8 bytes.
constexpr int kTileSame = 32;                          // same-colour pixels per tile row (a wave: 64 / kTileSame rows)
constexpr int kTileCols = 2 * kTileSame + 2 * kMaxHalf + 4;
struct LdsTaps {
    const float2* srow;
    const float2* scol;
    const uint8_t* sref;
    int r0, c0;
    PMP_DF float2 row(int ii) const { return srow[ii - r0]; }
    PMP_DF float2 col(int jj) const { return scol[jj - c0]; }
    PMP_DF uint32_t rowoff(int ii) const { return (uint32_t)((ii - r0) * kTileCols - c0); }
    PMP_DF uint32_t rowstep() const { return (uint32_t)kTileCols; }
    PMP_DF float refpx(uint32_t roff, int jj) const { return (float)sref[roff + (uint32_t)jj]; }
};

// Shifted moments of one view: a0 = sum w yd, a1 = sum w yd^2,
// a2 = sum w xc yd with yd = y - sh (sh = the centre tap's sample).
struct ViewMom {
    float a0, a1, a2, sh;
};

// Moments of views [v0, v0+G) (G <= 3 live views keep the per-view
// transforms in scalar registers; v0 is uniform).  The taps are walked from
// the centre tap tc (offsets (kc, kc), the middle of the row-major tap grid)
// to the last and on from the first to tc - 1, the same order for every
// view.  The centre tap's sample is the view's shift: it lies near the
// bilateral-weighted mean of the samples far more often than a corner tap,
// which keeps the moments' cancellation small.  It adds exact zeros to the
// three moments (yd = 0, w >= 0), so it is sampled before the loop and the
// loop starts at the tap after it.
template <int G, bool kZero, class Taps>
PMP_DF void view_group(const KfArgs& a, const Taps& tp, const float* sWc, const RefStat& rs, int i, int j, float px,
                       float py, float pz, int v0, ViewMom (&out)[G]) {
    const uint32_t* base[G];
#pragma unroll
    for (int g = 0; g < G; ++g) base[g] = a.vq[v0 + g];
    float acc0[G], acc1[G], acc2[G], shift[G];
    const float hu = in_vgpr(a.hu), hv = in_vgpr(a.hv);
    {
        const float2 rt = tp.row(i + a.kc);
        const float2 ct = tp.col(j + a.kc);
        float sbx, sby, sbz, as;
        tap_dir(rt.x * ct.x, -rt.y, rt.x * ct.y, px, py, pz, sbx, sby, sbz, as);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            shift[g] = view_sample<kZero>(a, hu, hv, base[g], v0 + g, sbx, sby, sbz, as);
            acc0[g] = 0.0f;
            acc1[g] = 0.0f;
            acc2[g] = 0.0f;
        }
    }
    // The other T-1 taps in that order, software-pipelined: view g's gather
    // for the next tap is issued right after view g's current tap is blended,
    // into the same registers, so each load has about one tap of work to
    // arrive.
    const int T = a.T;
    // tap index ti at (k1, k2) and roff = tp.rowoff(i + k1) + j, the
    // reference offset of its row (ti, k1, k2 are the same in every lane;
    // readfirstlane keeps them, and the row steps, in scalar registers
    // instead of per-lane copies); after the last tap the walk wraps to the
    // first, klast - first_k rows up
    const int first_k = __builtin_amdgcn_readfirstlane(-a.half);
    const uint32_t row_step = (uint32_t)a.stride * tp.rowstep();
    const uint32_t wrap_step = 0u - (uint32_t)(a.klast + a.half) * tp.rowstep();
    auto next = [&](int& ti, int& k1, int& k2, uint32_t& roff) {
        const int k2n = k2 + a.stride;
        const int k1n = k1 + a.stride;
        const bool wrap = k2n > a.half;
        const bool wrap_all = wrap && k1n > a.half;
        k2 = __builtin_amdgcn_readfirstlane(wrap ? first_k : k2n);
        k1 = __builtin_amdgcn_readfirstlane(wrap ? (wrap_all ? first_k : k1n) : k1);
        ti = __builtin_amdgcn_readfirstlane(ti + 1 == T ? 0 : ti + 1);
        roff += wrap ? (wrap_all ? wrap_step : row_step) : 0u;
    };
    // direction and reference weights of tap t at (k1, k2)
    auto tap = [&](int k1, int k2, uint32_t roff, int t, float& sbx, float& sby, float& sbz, float& as, float& w,
                   float& wxc) {
        const float2 rt = tp.row(i + k1);
        const float2 ct = tp.col(j + k2);
        tap_dir(rt.x * ct.x, -rt.y, rt.x * ct.y, px, py, pz, sbx, sby, sbz, as);
        const float xs = tp.refpx(roff, k2) - rs.Ic;
        w = sWc[(int)abs_(xs)] * a.Ws[t];
        wxc = w * (xs - rs.mx);
    };
    auto accumulate = [&](int g, float val, float w, float wxc) {
        const float yd = val - shift[g];
        const float tw = w * yd;
        acc0[g] = acc0[g] + tw;
        acc1[g] = fma_(tw, yd, acc1[g]);
        acc2[g] = fma_(wxc, yd, acc2[g]);
    };
    if (T > 1) {
        int ti = a.tc, k1 = a.kc, k2 = a.kc;
        uint32_t roff = tp.rowoff(i + a.kc) + (uint32_t)j;
        next(ti, k1, k2, roff);
        Pend<kZero> pd[G];
        float w, wxc;
        {
            float sbx, sby, sbz, as;
            tap(k1, k2, roff, ti, sbx, sby, sbz, as, w, wxc);
#pragma unroll
            for (int g = 0; g < G; ++g) view_gather<kZero>(a, hu, hv, base[g], v0 + g, sbx, sby, sbz, as, pd[g]);
        }
        // The scheduling barriers keep each view's blend and its next gather
        // in their own segment of the iteration; left alone the compiler
        // gathers all loads at the end of the iteration and waits for them
        // at the top of the next.
#pragma unroll 1
        for (int t = 1; t + 1 < T; ++t) {
            next(ti, k1, k2, roff);
            float sbx, sby, sbz, as, nw, nwxc;
            tap(k1, k2, roff, ti, sbx, sby, sbz, as, nw, nwxc);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int g = 0; g < G; ++g) {
                accumulate(g, blend<kZero>(pd[g]), w, wxc);
                view_gather<kZero>(a, hu, hv, base[g], v0 + g, sbx, sby, sbz, as, pd[g]);
                __builtin_amdgcn_sched_barrier(0);
            }
            w = nw;
            wxc = nwxc;
        }
#pragma unroll
        for (int g = 0; g < G; ++g) accumulate(g, blend<kZero>(pd[g]), w, wxc);
    }
#pragma unroll
    for (int g = 0; g < G; ++g) out[g] = ViewMom{acc0[g], acc1[g], acc2[g], shift[g]};
}

// Hands the moments of views [v0, v0+G) on: PMP_AGG_REF_FOLD pools them into
// the running state kept in LDS rows 0..4 of sv (fresh at v0 = 0); the other
// aggregations park each view's own NCC in row v.
template <int G>
PMP_DF void emit_views(const KfArgs& a, const RefStat& rs, int v0, const ViewMom (&m)[G], float* sv, int ld) {
    if (a.agg_mode == PMP_AGG_REF_FOLD) {
        Pool st{0.0f, 0.0f, 0.0f, 0.0f, -1.0f};
        if (v0 > 0) st = Pool{sv[0], sv[ld], sv[2 * ld], sv[3 * ld], sv[4 * ld]};
#pragma unroll
        for (int g = 0; g < G; ++g) pool_view(a, rs, st, a.rn[v0 + g], m[g].a0, m[g].a1, m[g].a2, m[g].sh);
        sv[0] = st.C;
        sv[ld] = st.V;
        sv[2 * ld] = st.mu;
        sv[3 * ld] = st.q;
        sv[4 * ld] = st.s;
    } else {
#pragma unroll
        for (int g = 0; g < G; ++g) sv[(v0 + g) * ld] = ncc(rs, a.var_floor, m[g].a0, m[g].a1, m[g].a2);
    }
}

// Score of plane (px, py, pz) at pixel (i, j); kZero selects PMP_EDGE_ZERO
// sampling.  a.N views (uniform) in groups of 3, 2 or 1; between groups the
// per-view scores (or the pooled state) are parked in LDS at sv[r * ld], r <
// kMaxViews, until they are aggregated.  NV must be 0 (run-time view count).
// tp: the source of the per-tap inputs (GlobalTaps, or LdsTaps for a tile).
template <int NV, bool kZero, class Taps>
PMP_DF float score_plane_t(const KfArgs& a, const Taps& tp, const float* sWc, const RefStat& rs, int i, int j,
                           float px, float py, float pz, float* sv, int ld) {
    static_assert(NV == 0, "the view count is a run-time value");
    const int N = a.N;
    for (int v0 = 0; v0 < N; v0 += 3) {
        const int G = (N - v0) < 3 ? (N - v0) : 3;
        if (G == 3) {
            ViewMom m[3];
            view_group<3, kZero>(a, tp, sWc, rs, i, j, px, py, pz, v0, m);
            emit_views<3>(a, rs, v0, m, sv, ld);
        } else if (G == 2) {
            ViewMom m[2];
            view_group<2, kZero>(a, tp, sWc, rs, i, j, px, py, pz, v0, m);
            emit_views<2>(a, rs, v0, m, sv, ld);
        } else {
            ViewMom m[1];
            view_group<1, kZero>(a, tp, sWc, rs, i, j, px, py, pz, v0, m);
            emit_views<1>(a, rs, v0, m, sv, ld);
        }
    }
    if (a.agg_mode == PMP_AGG_REF_FOLD) return sv[4 * ld];
    return aggregate_rt(a, sv, ld);
}

template <int NV, bool kZero>
PMP_DF float score_plane(const KfArgs& a, const float* sWc, const RefStat& rs, int i, int j, float px, float py,
                         float pz, float* sv = nullptr, int ld = 0) {
    return score_plane_t<NV, kZero>(a, global_taps(a), sWc, rs, i, j, px, py, pz, sv, ld);
}

// PlaneFromDepthNormal (PatchMatchCuda.cu:31-43)
PMP_DF void plane_from_depth_normal(float bx, float by, float bz, float dep, float nx, float ny, float nz,
                                    float& p0, float& p1, float& p2) {
    float qx = dep * bx, qy = dep * by, qz = dep * bz;
    float den = fma_(nz, qz, fma_(ny, qy, nx * qx));
    float r = rcp(den);
    p0 = nx * r;
    p1 = ny * r;
    p2 = nz * r;
}

// 1/fmaxf(1e-6, |b.p|) (PatchMatchCuda.cu:340-341, :402-403)
PMP_DF float depth_of_plane(float bx, float by, float bz, float p0, float p1, float p2) {
    float s = fma_(p2, bz, fma_(p1, by, p0 * bx));
    return rcp_pos(clampf(abs_(s), 1e-6f, 1e30f));
}

// Random hypothesis of one pixel (ComputeRandomDepthmap_/ComputeRandomPlanemap_,
// PatchMatchCuda.cu:210-244)
PMP_DF void random_plane(const KfArgs& a, int i, int j, float& p0, float& p1, float& p2, float& dep) {
    u32x4 r = philox((uint32_t)(i * a.W + j), 0xFFFFFFFFu, 0u, kStreamInit, a.key0, a.key1);
    float bx, by, bz;
    bearing(a, i, j, bx, by, bz);
    float nx, ny, nz;
    if (a.init_mode == PMP_INIT_INVDEPTH_HEMI) {
        float ud = u01_oc(r.x);
        float inv = fma_(ud, a.inv_hi - a.inv_lo, a.inv_lo);
        dep = rcp_pos(clampf(inv, 1e-30f, 1e30f));
        float zc = fma_(2.0f, u01_co(r.y), -1.0f);
        float rr = sqrt_(fma_(-zc, zc, 1.0f));
        float s, c;
        sincos2pi(u01_co(r.z), &s, &c);
        nx = rr * c;
        ny = rr * s;
        nz = zc;
        float dn = fma_(nz, bz, fma_(ny, by, nx * bx));
        if (dn > 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
    } else {
        dep = fma_(u01_oc(r.x), a.max_depth - a.min_depth, a.min_depth);
        nx = fma_(u01_oc(r.y), 2.0f, -1.0f);
        ny = fma_(u01_oc(r.z), 2.0f, -1.0f);
        nz = fma_(u01_oc(r.w), 2.0f, -1.0f);
    }
    plane_from_depth_normal(bx, by, bz, dep, nx, ny, nz, p0, p1, p2);
}

// Standard normals n of (pixel, iteration): counter (pixel, it, block, "RFNE").
template <int NMAX>
PMP_DF void refine_normals(const KfArgs& a, int i, int j, int it, int n, float (&z)[NMAX]) {
    const uint32_t pix = (uint32_t)(i * a.W + j);
#pragma unroll
    for (int b = 0; b < (NMAX + 3) / 4; ++b) {
        if (4 * b < n) {
            u32x4 r = philox(pix, (uint32_t)it, (uint32_t)b, kStreamRefine, a.key0, a.key1);
            float z0, z1, z2, z3;
            box_muller(r.x, r.y, &z0, &z1);
            box_muller(r.z, r.w, &z2, &z3);
            if (4 * b + 0 < NMAX) z[4 * b + 0] = z0;
            if (4 * b + 1 < NMAX) z[4 * b + 1] = z1;
            if (4 * b + 2 < NMAX) z[4 * b + 2] = z2;
            if (4 * b + 3 < NMAX) z[4 * b + 3] = z3;
        }
    }
}

}  // namespace pmp
