/*
 * or_core.c -- CPU oracle of the PatchMatch-Stereo-Panorama densification path.
 *
 * TEST INFRASTRUCTURE.  This is the checker the HIP path is compared against
 * (and, timed on the host, the CPU baseline).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * It restates the reference algorithm in plain C, pixel by pixel, in the
 * reference's own loop structure:
 *   bearings            PatchMatchCuda.cu:183-208
 *   random init         PatchMatchCuda.cu:210-244, :31-43 (plane from depth/normal)
 *   plane score         PatchMatchCuda.cu:46-180 (ComputeScorePlaneHypothesis)
 *   initial scoring     PatchMatchCuda.cu:255-270, grids :538-539
 *   red-black pass      PatchMatchCuda.cu:273-407, grid/launch order :578-592
 *   copy-back + gate    depthmap.cc:673-687, :731-760, :963-973
 *   PostProcess         depthmap.cc:1169-1208
 *   Cleaner             depthmap.cc:1320-1377 (+ :106-111, :143-161, :30-50)
 *   Pruner              depthmap.cc:1535-1696
 * with the canonical FP32 semantics of this project (one IEEE operation
 * sequence shared with the GPU, see or_math.h):
 *   - bearings from per-row (cos lat, sin lat) / per-column (sin lon, cos lon)
 *     tables computed in double and rounded once;
 *   - the two rigid transforms fused into M = R_n R_ref^T, c = t_n - M t_ref
 *     (double on the host, rounded once);
 *   - the bilateral weight exp(-dI^2 cf - |k|^2 xf) evaluated as the product
 *     of two float tables Wc[|dI|] * Ws[tap] (each exp in double);
 *   - shifted moments (same covariance and variances as the raw one-pass
 *     moments of :150-169, without their FP32 cancellation): reference taps
 *     centred on their weighted mean, neighbour samples shifted by the
 *     view's sample at the centre tap, the tap walk starting there;
 *   - explicit neighbour bilinear with longitude wrap / latitude clamp;
 *   - the reference fold (REF_FOLD) over moments pooled across views, as
 *     the device accumulates them (:57-62 zeroed once, outside the view loop
 *     :101): view v's NCC is that of views 0..v together, formed from the
 *     per-view shifted moments by running sums of the covariances and
 *     variances plus Welford's update of the spread of the view means;
 *   - NCC clamped to [-1, 1]; the fold's agreement test written without the
 *     division (|s - s_v| < agree |s_v|, identical at s_v = 0);
 *   - counter-based Philox4x32-10 random streams instead of clock-seeded
 *     XORWOW (deterministic; keyed on seed, counter = pixel/iteration/block).
 * Every function carries the reference file:line it follows.
 */
#define _GNU_SOURCE
#include <stdlib.h>
#include <stdio.h>
#include <math.h>
#include <float.h>
#include <stddef.h>
#include "or_api.h"
#include "or_math.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* derived per-keyframe quantities (host arithmetic, double -> float) */
/* ------------------------------------------------------------------ */
typedef struct {
    int W, H, N, T, half, stride, n1;
    int tc;   /* centre tap (the middle of the row-major tap grid): the shift's tap */
    int k1[OR_MAX_TAPS], k2[OR_MAX_TAPS];
    float Ws[OR_MAX_TAPS];
    float Wc[256];
    float* rowtab;   /* 2H: cos(lat_i), sin(lat_i) */
    float* coltab;   /* 2W: sin(lon_j), cos(lon_j) */
    float M[OR_MAX_VIEWS][9], c[OR_MAX_VIEWS][3];
    float ku, hu, kv, hv;
    float dr[OR_MAX_REFINE], nr[OR_MAX_REFINE];
    float inv_k, inv_T;
    float rn[OR_MAX_VIEWS];   /* 1/(v+1), rounded once */
    int topk;
    uint32_t key[2];
} derived_t;

/* bearing tables; follows ComputeBearingsEquirect_ (PatchMatchCuda.cu:191-198):
 * lon = 2 pi (j/W - 1/2), lat = -pi (i/H - 1/2), b = (cos lat sin lon, -sin lat, cos lat cos lon) */
EXPORT int or_tables(int W, int H, float* rowtab, float* coltab) {
    for (int i = 0; i < H; ++i) {
        double lat = -M_PI * ((double)i / (double)H - 0.5);
        rowtab[2 * i] = (float)cos(lat);
        rowtab[2 * i + 1] = (float)sin(lat);
    }
    for (int j = 0; j < W; ++j) {
        double lon = 2.0 * M_PI * ((double)j / (double)W - 0.5);
        coltab[2 * j] = (float)sin(lon);
        coltab[2 * j + 1] = (float)cos(lon);
    }
    return 0;
}

static inline void bearing_f(const derived_t* d, int i, int j, float* bx, float* by, float* bz) {
    float cl = d->rowtab[2 * i], sl = d->rowtab[2 * i + 1];
    float so = d->coltab[2 * j], co = d->coltab[2 * j + 1];
    *bx = cl * so;
    *by = -sl;
    *bz = cl * co;
}

/* fused neighbour transform; replaces the two-step R1_inv/t1_inv then R1n/t1n
 * of PatchMatchCuda.cu:124-130 with pose algebra of depthmap.cc:604-649 */
EXPORT int or_derive_poses(const double* ref_pose, int N, const double* view_poses, float* M, float* c) {
    const double* R0 = ref_pose; const double* t0 = ref_pose + 9;
    for (int n = 0; n < N; ++n) {
        const double* Rn = view_poses + 12 * n; const double* tn = Rn + 9;
        double Md[9];
        for (int r = 0; r < 3; ++r)
            for (int col = 0; col < 3; ++col) {
                /* (R_n R_0^T)[r][col] = sum_k Rn[r][k] R0[col][k] */
                double s = 0.0;
                for (int k = 0; k < 3; ++k) s += Rn[3 * r + k] * R0[3 * col + k];
                Md[3 * r + col] = s;
            }
        for (int r = 0; r < 3; ++r) {
            double s = tn[r];
            for (int k = 0; k < 3; ++k) s -= Md[3 * r + k] * t0[k];
            c[3 * n + r] = (float)s;
        }
        for (int k = 0; k < 9; ++k) M[9 * n + k] = (float)Md[k];
    }
    return 0;
}

static int derive(const or_keyframe* kf, const or_params* p, derived_t* d) {
    if (kf->N < 1 || kf->N > OR_MAX_VIEWS) return -1;
    if (p->patch_half < 3 || p->patch_stride < 1) return -1;
    d->W = kf->W; d->H = kf->H; d->N = kf->N;
    d->half = p->patch_half; d->stride = p->patch_stride;
    int T = 0;
    for (int a = -p->patch_half; a <= p->patch_half; a += p->patch_stride)
        for (int b = -p->patch_half; b <= p->patch_half; b += p->patch_stride) {
            if (T >= OR_MAX_TAPS) return -1;
            d->k1[T] = a; d->k2[T] = b; ++T;
        }
    d->T = T;
    {
        int n = 2 * d->half / d->stride + 1;   /* taps per row */
        d->tc = (n / 2) * n + n / 2;
    }
    double xf = 1.0 / (2.0 * (double)p->sigma_space * (double)p->sigma_space);
    double cf = 1.0 / (2.0 * (double)p->sigma_color * (double)p->sigma_color);
    for (int t = 0; t < T; ++t) d->Ws[t] = (float)exp(-(double)(d->k1[t] * d->k1[t] + d->k2[t] * d->k2[t]) * xf);
    for (int a = 0; a < 256; ++a) d->Wc[a] = (float)exp(-(double)(a * a) * cf);
    d->rowtab = (float*)malloc(sizeof(float) * 2 * d->H);
    d->coltab = (float*)malloc(sizeof(float) * 2 * d->W);
    or_tables(d->W, d->H, d->rowtab, d->coltab);
    or_derive_poses(kf->ref_pose, kf->N, kf->view_poses, &d->M[0][0], &d->c[0][0]);
    d->ku = (float)((double)d->W / (2.0 * M_PI));
    d->hu = (float)(0.5 * d->W);
    d->kv = (float)((double)d->H / M_PI);
    d->hv = (float)(0.5 * d->H);
    /* refinement ranges: PatchMatchCuda.cu:348-349, decays :393-394 (float *= double) */
    float dr = p->refine_depth_range, nr = p->refine_normal_range;
    for (int t = 0; t < OR_MAX_REFINE; ++t) {
        d->dr[t] = dr; d->nr[t] = nr;
        dr = (float)((double)dr * (double)p->refine_depth_decay);
        nr = (float)((double)nr * (double)p->refine_normal_decay);
    }
    int k = p->topk < 1 ? 1 : p->topk;
    if (k > d->N) k = d->N;
    d->topk = k;
    d->inv_k = (float)(1.0 / (double)k);
    d->inv_T = (float)(1.0 / (double)T);
    for (int v = 0; v < OR_MAX_VIEWS; ++v) d->rn[v] = (float)(1.0 / (double)(v + 1));
    d->key[0] = (uint32_t)(p->seed & 0xFFFFFFFFu);
    d->key[1] = (uint32_t)(p->seed >> 32);
    return 0;
}

static void release(derived_t* d) { free(d->rowtab); free(d->coltab); }

/* ------------------------------------------------------------------ */
/* coverage (PatchMatchCuda.cu:262,283 border; grids :465-470, :538-539, :578-581) */
/* ------------------------------------------------------------------ */
static inline int interior(const derived_t* d, int i, int j) {
    int ph = d->half;
    return !(i < ph + 1 || i > d->H - ph - 1 || j < ph + 1 || j > d->W - ph - 1);
}
static inline int in_init_grid(const derived_t* d, const or_params* p, int i, int j) {
    if (p->coverage_mode == OR_COVER_FULL) return 1;
    return i < (d->H / 16) * 16 && j < (d->W / 16) * 16;
}
static inline int in_score_grid(const derived_t* d, const or_params* p, int i, int j) {
    if (!interior(d, i, j)) return 0;
    if (p->coverage_mode == OR_COVER_FULL) return 1;
    return i < (d->H / 32) * 32 && j < (d->W / 32) * 32;
}
static inline int in_prop_grid(const derived_t* d, const or_params* p, int i, int j) {
    if (!interior(d, i, j)) return 0;
    if (p->coverage_mode == OR_COVER_FULL) return 1;
    int J = 32 * ((d->W / 32) / 2);
    return i < (d->H / 32) * 32 && (j + (j & 1)) < 2 * J;
}

/* ------------------------------------------------------------------ */
/* sampling of a neighbour image at (u, v); texel k sits at integer k   */
/* (lookup at (u+0.5, v+0.5) in PatchMatchCuda.cu:146 with point centres) */
/* ------------------------------------------------------------------ */
static inline float sample_f(const uint8_t* img, int W, int H, int edge, float u, float v) {
    float b0, b1, b2, b3, fx, fy;
    if (edge == OR_EDGE_ZERO) {
        /* depthmap.cc:36-38: outside [0, W-1) x [0, H-1) -> 0 */
        if (!(u >= 0.0f && u < (float)(W - 1) && v >= 0.0f && v < (float)(H - 1))) return 0.0f;
        float x0f = floorf(u), y0f = floorf(v);
        int x0 = (int)x0f, y0 = (int)y0f;
        fx = u - x0f; fy = v - y0f;
        const uint8_t* r0 = img + (size_t)y0 * W;
        const uint8_t* r1 = r0 + W;
        b0 = r0[x0]; b1 = r0[x0 + 1]; b2 = r1[x0]; b3 = r1[x0 + 1];
    } else {
        float uu = or_clamp(u, 0.0f, edge == OR_EDGE_WRAP ? (float)W : (float)(W - 1));
        float vv = or_clamp(v, 0.0f, (float)(H - 1));
        float x0f = floorf(uu), y0f = floorf(vv);
        int x0 = (int)x0f, y0 = (int)y0f;
        fx = uu - x0f; fy = vv - y0f;
        if (x0 >= W) x0 -= W;
        int x1 = x0 + 1;
        if (x1 >= W) x1 = (edge == OR_EDGE_WRAP) ? 0 : W - 1;
        int y1 = (y0 + 1 < H) ? y0 + 1 : H - 1;
        const uint8_t* r0 = img + (size_t)y0 * W;
        const uint8_t* r1 = img + (size_t)y1 * W;
        b0 = r0[x0]; b1 = r0[x1]; b2 = r1[x0]; b3 = r1[x1];
    }
    float top = fmaf(fx, b1 - b0, b0);
    float bot = fmaf(fx, b3 - b2, b2);
    return fmaf(fy, bot - top, top);
}

/* equirect projection of a camera-frame point; PatchMatchCuda.cu:132-144
 * (its latitude asin(y/|X|) is the angle atan2(y, sqrt(x^2+z^2)) of or_lonlat) */
static inline void project_f(const derived_t* d, float x, float y, float z, float* u, float* v) {
    float lon, lat;
    or_lonlat(x, y, z, &lon, &lat);
    *u = fmaf(lon, d->ku, d->hu);
    *v = fmaf(lat, d->kv, d->hv);
}

/* Ray-plane intersection of tap bearing b with plane p (PatchMatchCuda.cu:117-122,
 * |b.p| clamped to >= 1e-6 like depthmap.cc:167-173) moved into view v.  The
 * camera-frame point is b / s, s = b.p; its view-frame position M b / s + c
 * has the direction of sigma M b + c |s| (sigma = sign bit of s), which is all
 * the equirect projection uses.  tap_dir_f folds sigma into b and returns
 * max(|s|, 1e-6); view_dir_f forms the direction for view v. */
static inline void tap_dir_f(float bx, float by, float bz, float px, float py, float pz, float* sb, float* as) {
    float s = fmaf(pz, bz, fmaf(py, by, px * bx));
    uint32_t sg = or_f2u(s) & 0x80000000u;
    sb[0] = or_u2f(or_f2u(bx) ^ sg);
    sb[1] = or_u2f(or_f2u(by) ^ sg);
    sb[2] = or_u2f(or_f2u(bz) ^ sg);
    *as = or_max(fabsf(s), 1e-6f);
}

static inline void view_dir_f(const float* M, const float* c, const float* sb, float as, float* x, float* y,
                              float* z) {
    *x = fmaf(c[0], as, fmaf(M[0], sb[0], fmaf(M[1], sb[1], M[2] * sb[2])));
    *y = fmaf(c[1], as, fmaf(M[3], sb[0], fmaf(M[4], sb[1], M[5] * sb[2])));
    *z = fmaf(c[2], as, fmaf(M[6], sb[0], fmaf(M[7], sb[1], M[8] * sb[2])));
}

/* ------------------------------------------------------------------ */
/* plane-hypothesis score; ComputeScorePlaneHypothesis PatchMatchCuda.cu:46-180 */
/* ------------------------------------------------------------------ */
typedef struct { float Ic, mx, inv, mxc, vx; } refstat_t;

/* reference-side moments of the patch: the plane-independent part of :148-169.
 * Two passes: weighted mean mx of x' = I - Ic, then the centred taps
 * xc = x' - mx give mxc = E[xc] (~0) and vx = E[xc^2]. */
static inline refstat_t ref_stats(const derived_t* d, const uint8_t* ref, int i, int j) {
    refstat_t s;
    int W = d->W;
    s.Ic = (float)ref[(size_t)i * W + j];
    float Sw = 0.0f, Swx = 0.0f;
    for (int t = 0; t < d->T; ++t) {
        float xs = (float)ref[(size_t)(i + d->k1[t]) * W + (j + d->k2[t])] - s.Ic;
        float w = d->Wc[(int)fabsf(xs)] * d->Ws[t];
        Sw = Sw + w;
        Swx = fmaf(w, xs, Swx);
    }
    s.inv = or_rcp_pos(Sw);
    s.mx = Swx * s.inv;
    float Sc = 0.0f, Scc = 0.0f;
    for (int t = 0; t < d->T; ++t) {
        float xs = (float)ref[(size_t)(i + d->k1[t]) * W + (j + d->k2[t])] - s.Ic;
        float w = d->Wc[(int)fabsf(xs)] * d->Ws[t];
        float xc = xs - s.mx;
        float wxc = w * xc;
        Sc = Sc + wxc;
        Scc = fmaf(wxc, xc, Scc);
    }
    s.mxc = Sc * s.inv;
    s.vx = fmaf(-s.mxc, s.mxc, Scc * s.inv);
    return s;
}

/* NCC of one view from the shifted moments: cov = E[xc y] - E[xc] E[y],
 * var_y = E[y^2] - E[y]^2 (y shifted by the view's centre-tap sample) */
static inline float ncc_f(const refstat_t* rs, float floor_, float Swy, float Swyy, float Swxy) {
    float my = Swy * rs->inv;
    float vy = fmaf(-my, my, Swyy * rs->inv);
    float cov = fmaf(-rs->mxc, my, Swxy * rs->inv);
    if (rs->vx < floor_ || vy < floor_) return -1.0f;
    float s = cov * or_rsqrt_pos(rs->vx * vy);
    return or_clamp(s, -1.0f, 1.0f);
}

/* the fold of PatchMatchCuda.cu:171: s <- (|(s - s2)/s2| < agree) ? (s + s2)/2 : max(s, s2) */
static inline float fold_f(float s, float s2, float agree) {
    if (fabsf(s - s2) < agree * fabsf(s2)) return 0.5f * (s + s2);
    return (s2 > s) ? s2 : s;
}

/* Moments pooled across views (REF_FOLD).  PatchMatchCuda.cu:57-62 zero the
 * six sums once and the view loop (:101-157) keeps adding to them, so the
 * NCC of view v at :161-169 is that of the taps of views 0..v together: the
 * reference side repeats, so its mean and variance stay those of one view;
 * the pooled covariance is the mean of the per-view covariances; the pooled
 * variance of y is the mean of the per-view variances plus the spread of the
 * per-view means.  C and V sum the covariances and variances, mu and q are
 * Welford's running mean and sum of squared deviations of the view means. */
typedef struct { float C, V, mu, q, s; } pool_t;

static inline void pool_init(pool_t* st) { st->C = 0.0f; st->V = 0.0f; st->mu = 0.0f; st->q = 0.0f; st->s = -1.0f; }

/* adds view v (shifted moments acc0..2 about its shift), returns the pooled NCC
 * of views 0..v and folds it into st->s */
static inline float pool_view(pool_t* st, const refstat_t* rs, float floor_, float agree, float rn, float acc0,
                              float acc1, float acc2, float shift) {
    float my = acc0 * rs->inv;
    float vy = fmaf(-my, my, acc1 * rs->inv);
    float cov = fmaf(-rs->mxc, my, acc2 * rs->inv);
    float m = shift + my;
    st->C = st->C + cov;
    st->V = st->V + vy;
    float dm = m - st->mu;
    st->mu = fmaf(dm, rn, st->mu);
    st->q = fmaf(dm, m - st->mu, st->q);
    float cp = st->C * rn;
    float vp = (st->V + st->q) * rn;
    float s2 = -1.0f;
    if (!(rs->vx < floor_ || vp < floor_)) s2 = or_clamp(cp * or_rsqrt_pos(rs->vx * vp), -1.0f, 1.0f);
    st->s = fold_f(st->s, s2, agree);
    return s2;
}

/* view aggregation of independent per-view scores: fold PatchMatchCuda.cu:170-171
 * (REF_FOLD given the pooled scores, FOLD_INDEP) / max depthmap.cc:1110-1117 / top-k mean */
static inline float aggregate_f(const derived_t* d, const or_params* p, const float* sv) {
    int N = d->N;
    if (p->agg_mode == OR_AGG_MAX) {
        float s = -1.0f;
        for (int v = 0; v < N; ++v) s = (sv[v] > s) ? sv[v] : s;
        return s;
    }
    if (p->agg_mode == OR_AGG_TOPK_MEAN) {
        float tmp[OR_MAX_VIEWS];
        for (int v = 0; v < N; ++v) tmp[v] = sv[v];
        /* selection of the k largest, in descending order */
        float sum = 0.0f;
        for (int a = 0; a < d->topk; ++a) {
            int best = a;
            for (int b = a + 1; b < N; ++b) if (tmp[b] > tmp[best]) best = b;
            float t = tmp[a]; tmp[a] = tmp[best]; tmp[best] = t;
            sum = (a == 0) ? tmp[0] : sum + tmp[a];
        }
        return sum * d->inv_k;
    }
    float s = -1.0f;
    for (int v = 0; v < N; ++v) s = fold_f(s, sv[v], p->fold_agree);
    return s;
}

static float score_plane_f(const derived_t* d, const or_params* p, const or_keyframe* kf,
                           const refstat_t* rs, int i, int j, float px, float py, float pz,
                           float* view_scores /* nullable */) {
    int N = d->N, W = d->W, H = d->H;
    float acc0[OR_MAX_VIEWS], acc1[OR_MAX_VIEWS], acc2[OR_MAX_VIEWS];
    for (int v = 0; v < N; ++v) { acc0[v] = 0.0f; acc1[v] = 0.0f; acc2[v] = 0.0f; }
    float shift[OR_MAX_VIEWS];
    const uint8_t* ref = kf->ref;
    /* taps from the centre tap tc to the last, then from the first to tc - 1;
     * the centre tap's sample is each view's shift (it lies near the weighted
     * mean far more often than a corner tap: smaller cancellation) */
    for (int q = 0; q < d->T; ++q) {
        int t = (q + d->tc) % d->T;
        int ii = i + d->k1[t], jj = j + d->k2[t];
        float bx, by, bz, sb[3], as;
        bearing_f(d, ii, jj, &bx, &by, &bz);
        tap_dir_f(bx, by, bz, px, py, pz, sb, &as);
        float xs = (float)ref[(size_t)ii * W + jj] - rs->Ic;
        float w = d->Wc[(int)fabsf(xs)] * d->Ws[t];
        float wxc = w * (xs - rs->mx);
        for (int v = 0; v < N; ++v) {
            float x, y, z;
            view_dir_f(d->M[v], d->c[v], sb, as, &x, &y, &z);
            float u, vv;
            project_f(d, x, y, z, &u, &vv);
            float val = sample_f(kf->views[v], W, H, p->edge_mode, u, vv);
            if (q == 0) shift[v] = val;
            float yd = val - shift[v];
            float tw = w * yd;
            acc0[v] = acc0[v] + tw;
            acc1[v] = fmaf(tw, yd, acc1[v]);
            acc2[v] = fmaf(wxc, yd, acc2[v]);
        }
    }
    float sv[OR_MAX_VIEWS];
    if (p->agg_mode == OR_AGG_REF_FOLD) {
        pool_t st;
        pool_init(&st);
        for (int v = 0; v < N; ++v)
            sv[v] = pool_view(&st, rs, p->var_floor, p->fold_agree, d->rn[v], acc0[v], acc1[v], acc2[v], shift[v]);
        if (view_scores) for (int v = 0; v < N; ++v) view_scores[v] = sv[v];
        return st.s;
    }
    for (int v = 0; v < N; ++v) sv[v] = ncc_f(rs, p->var_floor, acc0[v], acc1[v], acc2[v]);
    if (view_scores) for (int v = 0; v < N; ++v) view_scores[v] = sv[v];
    return aggregate_f(d, p, sv);
}

/* ------------------------------------------------------------------ */
/* plane from depth and normal, PlaneFromDepthNormal PatchMatchCuda.cu:31-43 */
/* ------------------------------------------------------------------ */
static inline void plane_from_depth_normal(float bx, float by, float bz, float dep,
                                           float nx, float ny, float nz, float* p) {
    float qx = dep * bx, qy = dep * by, qz = dep * bz;
    float den = fmaf(nz, qz, fmaf(ny, qy, nx * qx));
    float r = or_rcp(den);
    p[0] = nx * r; p[1] = ny * r; p[2] = nz * r;
}

static inline float depth_of_plane(float bx, float by, float bz, const float* p) {
    /* 1/fmaxf(1e-6, |b.p|), PatchMatchCuda.cu:340-341, :402-403 */
    float s = fmaf(p[2], bz, fmaf(p[1], by, p[0] * bx));
    return or_rcp_pos(or_clamp(fabsf(s), 1e-6f, 1e30f));
}

/* random plane of one pixel (ComputeRandomDepthmap_/ComputeRandomPlanemap_, :210-244) */
static void random_plane(const derived_t* d, const or_params* p, int i, int j, float* plane, float* depth) {
    uint32_t ctr[4] = { (uint32_t)(i * d->W + j), 0xFFFFFFFFu, 0u, OR_STREAM_INIT };
    uint32_t r[4];
    or_philox(ctr, d->key, r);
    float bx, by, bz;
    bearing_f(d, i, j, &bx, &by, &bz);
    float dep, nx, ny, nz;
    if (p->init_mode == OR_INIT_INVDEPTH_HEMI) {
        float ud = or_u01_oc(r[0]);
        float inv = fmaf(ud, p->init_inv_depth_max - p->init_inv_depth_min, p->init_inv_depth_min);
        dep = or_rcp_pos(or_clamp(inv, 1e-30f, 1e30f));
        float zc = fmaf(2.0f, or_u01_co(r[1]), -1.0f);
        float rr = or_sqrt(fmaf(-zc, zc, 1.0f));
        float s, c;
        or_sincos2pi(or_u01_co(r[2]), &s, &c);
        nx = rr * c; ny = rr * s; nz = zc;
        float dn = fmaf(nz, bz, fmaf(ny, by, nx * bx));
        if (dn > 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
    } else {
        /* curand_between: u * (max - min) + min, u in (0, 1] */
        dep = fmaf(or_u01_oc(r[0]), p->max_depth - p->min_depth, p->min_depth);
        nx = fmaf(or_u01_oc(r[1]), 2.0f, -1.0f);
        ny = fmaf(or_u01_oc(r[2]), 2.0f, -1.0f);
        nz = fmaf(or_u01_oc(r[3]), 2.0f, -1.0f);
    }
    plane_from_depth_normal(bx, by, bz, dep, nx, ny, nz, plane);
    *depth = dep;
}

/* 18 (up to 3*OR_MAX_REFINE) standard normals of pixel/iteration */
static void refine_normals(const derived_t* d, int i, int j, int it, int n, float* z) {
    int words = 2 * ((n + 1) / 2);
    uint32_t buf[4 * ((3 * OR_MAX_REFINE + 3) / 4 + 1)];
    int blocks = (words + 3) / 4;
    for (int b = 0; b < blocks; ++b) {
        uint32_t ctr[4] = { (uint32_t)(i * d->W + j), (uint32_t)it, (uint32_t)b, OR_STREAM_REFINE };
        or_philox(ctr, d->key, buf + 4 * b);
    }
    for (int k = 0; k < n; k += 2) {
        float z0, z1;
        or_box_muller(buf[k], buf[k + 1], &z0, &z1);
        z[k] = z0;
        if (k + 1 < n) z[k + 1] = z1;
    }
}

static const uint32_t NAN_BITS = 0xFFFFFFFFu; /* cudaMemset(score, -1.0f -> 0xFF bytes), depthmap.cc:657 */

/* ------------------------------------------------------------------ */
/* exported stages                                                     */
/* ------------------------------------------------------------------ */

/* RandomInitialization (PatchMatchCuda.cu:656-663) after the memsets of depthmap.cc:657-659 */
EXPORT int or_init(const or_keyframe* kf, const or_params* p, float* plane4, float* depth, float* score) {
    derived_t d;
    if (derive(kf, p, &d)) return -1;
    int W = d.W, H = d.H;
    float nanv = or_u2f(NAN_BITS);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            size_t k = (size_t)i * W + j;
            float pl[3] = {0.0f, 0.0f, 0.0f}, dep = 0.0f;
            if (in_init_grid(&d, p, i, j)) random_plane(&d, p, i, j, pl, &dep);
            plane4[4 * k] = pl[0]; plane4[4 * k + 1] = pl[1]; plane4[4 * k + 2] = pl[2]; plane4[4 * k + 3] = 0.0f;
            depth[k] = dep;
            score[k] = nanv;
        }
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            if (!in_score_grid(&d, p, i, j)) continue;
            size_t k = (size_t)i * W + j;
            refstat_t rs = ref_stats(&d, kf->ref, i, j);
            score[k] = score_plane_f(&d, p, kf, &rs, i, j, plane4[4 * k], plane4[4 * k + 1], plane4[4 * k + 2], NULL);
        }
    release(&d);
    return 0;
}

/* score of a fixed plane field at every interior pixel (ScorePlaneDepth_, :255-270);
 * non-interior pixels get NaN; view_scores (nullable) gets the N per-view NCC
 * (REF_FOLD: view v's NCC over the moments pooled over views 0..v) */
EXPORT int or_score_field(const or_keyframe* kf, const or_params* p, const float* plane4, float* score,
                          float* view_scores) {
    derived_t d;
    if (derive(kf, p, &d)) return -1;
    int W = d.W, H = d.H, N = d.N;
    float nanv = or_u2f(NAN_BITS);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            size_t k = (size_t)i * W + j;
            if (!interior(&d, i, j)) {
                score[k] = nanv;
                if (view_scores) for (int v = 0; v < N; ++v) view_scores[k * N + v] = nanv;
                continue;
            }
            refstat_t rs = ref_stats(&d, kf->ref, i, j);
            score[k] = score_plane_f(&d, p, kf, &rs, i, j, plane4[4 * k], plane4[4 * k + 1], plane4[4 * k + 2],
                                     view_scores ? view_scores + k * N : NULL);
        }
    release(&d);
    return 0;
}

/* one pixel of PatchMatchRedBlackPass_ (PatchMatchCuda.cu:273-407) */
static void update_pixel(const derived_t* d, const or_params* p, const or_keyframe* kf, int it,
                         float* plane4, float* depth, float* score, int i, int j, int64_t* evals) {
    static const int adj[8][2] = { {0, 1}, {1, 0}, {0, -1}, {-1, 0}, {0, 3}, {3, 0}, {0, -3}, {-3, 0} };
    int W = d->W;
    size_t k = (size_t)i * W + j;
    refstat_t rs = ref_stats(d, kf->ref, i, j);
    float sc = score[k];
    float ph[3] = { plane4[4 * k], plane4[4 * k + 1], plane4[4 * k + 2] };
    float bx, by, bz;
    bearing_f(d, i, j, &bx, &by, &bz);
    float d0 = depth[k];
    float inv_d0 = or_rcp(d0);
    int64_t ev = 0;
    /* 1. propagation :311-338 */
    for (int a = 0; a < 8; ++a) {
        size_t kn = (size_t)(i + adj[a][0]) * W + (j + adj[a][1]);
        float pn[3] = { plane4[4 * kn], plane4[4 * kn + 1], plane4[4 * kn + 2] };
        float s = fmaf(pn[2], bz, fmaf(pn[1], by, pn[0] * bx));
        float dn = or_rcp_pos(or_clamp(fabsf(s), 1e-6f, 1e30f));
        float rel = (d0 - dn) * inv_d0;
        float fd = or_exp((rel * rel) * -p->depth_penalty);
        float prod = fmaf(ph[2], pn[2], fmaf(ph[1], pn[1], ph[0] * pn[0]));
        float ex = ph[0] - pn[0], ey = ph[1] - pn[1], ez = ph[2] - pn[2];
        float nrm = or_sqrt(fmaf(ez, ez, fmaf(ey, ey, ex * ex)));
        /* acosf(prod/norm) is NaN (candidate rejected) unless |prod| <= norm, norm > 0 */
        if (!(nrm > 0.0f && fabsf(prod) <= nrm)) continue;
        float ratio = or_clamp(prod * or_rcp_pos(or_clamp(nrm, 1e-30f, 1e30f)), -1.0f, 1.0f);
        float ac = or_acos(ratio);
        float fp = or_exp((ac * ac) * -p->plane_penalty);
        float f = (1.0f - fd) * (1.0f - fp);
        /* |score_hyp| <= 1, so score_hyp * f <= f: no acceptance unless f > score */
        if (!(f > sc)) continue;
        float sh = score_plane_f(d, p, kf, &rs, i, j, pn[0], pn[1], pn[2], NULL);
        ++ev;
        sh = sh * f;
        if (sh > sc) { ph[0] = pn[0]; ph[1] = pn[1]; ph[2] = pn[2]; sc = sh; }
    }
    /* 2. refinement :340-395 */
    float dcur = depth_of_plane(bx, by, bz, ph);
    float cur[3] = { ph[0], ph[1], ph[2] };
    float z[3 * OR_MAX_REFINE];
    int nref = p->n_refine;
    refine_normals(d, i, j, it, 3 * nref, z);
    for (int t = 0; t < nref; ++t) {
        float ra = z[3 * t], rb = z[3 * t + 1], rc = z[3 * t + 2];
        float dh = dcur * or_exp(d->dr[t] * ra);
        float nx = fmaf(d->nr[t], ra, cur[0]);
        float ny = fmaf(d->nr[t], rb, cur[1]);
        float nz = fmaf(d->nr[t], rc, cur[2]);
        float hyp[3];
        plane_from_depth_normal(bx, by, bz, dh, nx, ny, nz, hyp);
        float sh = score_plane_f(d, p, kf, &rs, i, j, hyp[0], hyp[1], hyp[2], NULL);
        ++ev;
        if (sh > sc) { cur[0] = hyp[0]; cur[1] = hyp[1]; cur[2] = hyp[2]; sc = sh; dcur = dh; }
    }
    score[k] = sc;
    plane4[4 * k] = cur[0]; plane4[4 * k + 1] = cur[1]; plane4[4 * k + 2] = cur[2];
    depth[k] = depth_of_plane(bx, by, bz, cur);
    *evals += ev;
}

/* iterations [it_begin, it_end) of PatchMatchRedBlackPass (PatchMatchCuda.cu:585-592):
 * pass 0 updates (i+j) even, pass 1 (i+j) odd.  evals_out (nullable) counts scored hypotheses. */
EXPORT int or_propagate(const or_keyframe* kf, const or_params* p, float* plane4, float* depth, float* score,
                        int it_begin, int it_end, int64_t* evals_out) {
    derived_t d;
    if (derive(kf, p, &d)) return -1;
    if (p->n_refine < 0 || p->n_refine > OR_MAX_REFINE) { release(&d); return -1; }
    int W = d.W, H = d.H;
    int64_t evals = 0;
    for (int it = it_begin; it < it_end; ++it)
        for (int colour = 0; colour < 2; ++colour) {
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : evals)
            for (int i = 0; i < H; ++i)
                for (int j = ((i + colour) & 1); j < W; j += 2) {
                    if (!in_prop_grid(&d, p, i, j)) continue;
                    update_pixel(&d, p, kf, it, plane4, depth, score, i, j, &evals);
                }
        }
    if (evals_out) *evals_out = evals;
    release(&d);
    return 0;
}

/* restricted variant for the CPU baseline: only rows [row_begin, row_end) */
EXPORT int or_propagate_rows(const or_keyframe* kf, const or_params* p, float* plane4, float* depth, float* score,
                             int it, int colour, int row_begin, int row_end, int64_t* evals_out) {
    derived_t d;
    if (derive(kf, p, &d)) return -1;
    int W = d.W;
    int64_t evals = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : evals)
    for (int i = row_begin; i < row_end; ++i)
        for (int j = ((i + colour) & 1); j < W; j += 2) {
            if (!in_prop_grid(&d, p, i, j)) continue;
            update_pixel(&d, p, kf, it, plane4, depth, score, i, j, &evals);
        }
    if (evals_out) *evals_out = evals;
    release(&d);
    return 0;
}

/* ------------------------------------------------------------------ */
/* median 3x3, border replicate (cv::medianBlur(src, dst, 3) for CV_32F) */
/* ------------------------------------------------------------------ */
static void median3(const float* src, float* dst, int W, int H) {
#pragma omp parallel for
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            float v[9];
            int n = 0;
            for (int a = -1; a <= 1; ++a)
                for (int b = -1; b <= 1; ++b) {
                    int ii = i + a < 0 ? 0 : (i + a >= H ? H - 1 : i + a);
                    int jj = j + b < 0 ? 0 : (j + b >= W ? W - 1 : j + b);
                    v[n++] = src[(size_t)ii * W + jj];
                }
            for (int a = 1; a < 9; ++a) {           /* insertion sort */
                float x = v[a]; int b = a - 1;
                while (b >= 0 && v[b] > x) { v[b + 1] = v[b]; --b; }
                v[b + 1] = x;
            }
            dst[(size_t)i * W + j] = v[4];
        }
}

/* copy-back window, low-texture gate and PostProcess of DepthmapEstimator::ComputePatchMatch
 * (depthmap.cc:673-687 init copy, :731-760 window, :767 -> :1169-1208).
 * init_* are the maps right after RandomInitialization; fin_* after the red-black passes.
 * Outputs: depth (post-processed), plane3, score, nghbr. */
EXPORT int or_finalize(const or_keyframe* kf, const or_params* p, const uint8_t* mask,
                       const float* init_plane4, const float* init_depth, const float* init_score,
                       const float* fin_plane4, const float* fin_depth, const float* fin_score,
                       float* out_depth, float* out_plane3, float* out_score, int32_t* out_nghbr) {
    derived_t d;
    if (derive(kf, p, &d)) return -1;
    int W = d.W, H = d.H;
    int ps = d.half + 2;                           /* patch_size_ with patchhalf = patch_size_-2 */
    int lo = ps - 2, hi_r = H - ps - 2, hi_c = W - ps - 2;
    float min_var = p->min_patch_sd * p->min_patch_sd;
#pragma omp parallel for
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            size_t k = (size_t)i * W + j;
            float dep = 0.0f, sc = -1.0f, pl[3] = {0.0f, 0.0f, 0.0f};
            int ng = 0;
            if (mask == NULL || mask[k]) {
                if (i >= lo && i < hi_r && j >= lo && j < hi_c) {
                    int low = 0;
                    if (p->min_patch_sd > 0.0f) {
                        /* PatchVariance (:963-973) over the T tap samples */
                        float sum = 0.0f;
                        for (int t = 0; t < d.T; ++t) sum = sum + (float)kf->ref[(size_t)(i + d.k1[t]) * W + (j + d.k2[t])];
                        float mean = sum * d.inv_T;
                        float s2 = 0.0f;
                        for (int t = 0; t < d.T; ++t) {
                            float e = (float)kf->ref[(size_t)(i + d.k1[t]) * W + (j + d.k2[t])] - mean;
                            s2 = fmaf(e, e, s2);
                        }
                        low = (s2 * d.inv_T) < min_var;
                    }
                    if (!low) {
                        dep = fin_depth[k]; sc = fin_score[k];
                        pl[0] = fin_plane4[4 * k]; pl[1] = fin_plane4[4 * k + 1]; pl[2] = fin_plane4[4 * k + 2];
                    } else {
                        sc = 0.0f;
                    }
                    ng = 1;
                } else {
                    dep = init_depth[k]; sc = init_score[k];
                    pl[0] = init_plane4[4 * k]; pl[1] = init_plane4[4 * k + 1]; pl[2] = init_plane4[4 * k + 2];
                }
            }
            out_depth[k] = dep; out_score[k] = sc; out_nghbr[k] = ng;
            out_plane3[3 * k] = pl[0]; out_plane3[3 * k + 1] = pl[1]; out_plane3[3 * k + 2] = pl[2];
        }
    if (p->postprocess) {
        float* m1 = (float*)malloc(sizeof(float) * (size_t)W * H);
        float* m2 = (float*)malloc(sizeof(float) * (size_t)W * H);
        median3(out_depth, m1, W, H);
        median3(m1, m2, W, H);
#pragma omp parallel for
        for (int i = 0; i < H; ++i)
            for (int j = 0; j < W; ++j) {
                size_t k = (size_t)i * W + j;
                float dd = out_depth[k], m = m2[k], sc = out_score[k];
                /* the float quotient against the double threshold (:1190: fabs of a
                 * float is float there, 0.05 a double literal) */
                if (dd == 0.0f || (double)(fabsf(dd - m) / dd) > p->median_rel_thr ||
                    dd < p->min_depth || dd > p->max_depth || sc <= p->min_score) {
                    out_depth[k] = 0.0f; out_score[k] = -1.0f;
                } else {
                    out_depth[k] = m;
                }
            }
        free(m1); free(m2);
    }
    release(&d);
    return 0;
}

/* the whole estimator for one keyframe (DepthmapEstimator::ComputePatchMatch without
 * the host pre-init, depthmap.cc:541-806); raw_* (nullable) receive the maps after
 * the last red-black pass */
EXPORT int or_densify(const or_keyframe* kf, const or_params* p, const uint8_t* mask,
                      float* out_depth, float* out_plane3, float* out_score, int32_t* out_nghbr,
                      float* raw_plane4, float* raw_depth, float* raw_score) {
    size_t P = (size_t)kf->W * kf->H;
    float* ip = (float*)malloc(sizeof(float) * 4 * P);
    float* id = (float*)malloc(sizeof(float) * P);
    float* is = (float*)malloc(sizeof(float) * P);
    float* fp = (float*)malloc(sizeof(float) * 4 * P);
    float* fd = (float*)malloc(sizeof(float) * P);
    float* fs = (float*)malloc(sizeof(float) * P);
    int rc = or_init(kf, p, ip, id, is);
    if (!rc) {
        for (size_t k = 0; k < 4 * P; ++k) fp[k] = ip[k];
        for (size_t k = 0; k < P; ++k) { fd[k] = id[k]; fs[k] = is[k]; }
        rc = or_propagate(kf, p, fp, fd, fs, 0, p->iterations, NULL);
    }
    if (!rc) rc = or_finalize(kf, p, mask, ip, id, is, fp, fd, fs, out_depth, out_plane3, out_score, out_nghbr);
    if (!rc && raw_plane4) for (size_t k = 0; k < 4 * P; ++k) raw_plane4[k] = fp[k];
    if (!rc && raw_depth) for (size_t k = 0; k < P; ++k) raw_depth[k] = fd[k];
    if (!rc && raw_score) for (size_t k = 0; k < P; ++k) raw_score[k] = fs[k];
    free(ip); free(id); free(is); free(fp); free(fd); free(fs);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Cleaner and Pruner (double geometry as in the reference)           */
/* ------------------------------------------------------------------ */
typedef struct { float* rowtab; float* coltab; float ku, hu, kv, hv; int W, H; } geo_t;

static void geo_init(geo_t* g, int W, int H) {
    g->W = W; g->H = H;
    g->rowtab = (float*)malloc(sizeof(float) * 2 * H);
    g->coltab = (float*)malloc(sizeof(float) * 2 * W);
    or_tables(W, H, g->rowtab, g->coltab);
    g->ku = (float)((double)W / (2.0 * M_PI)); g->hu = (float)(0.5 * W);
    g->kv = (float)((double)H / M_PI); g->hv = (float)(0.5 * H);
}
static void geo_free(geo_t* g) { free(g->rowtab); free(g->coltab); }

/* Backproject_equirect (depthmap.cc:106-111): R^T (d b - t), double */
static inline void backproject(const geo_t* g, const double* pose, int i, int j, float dep, double* X) {
    double bx = (double)(g->rowtab[2 * i] * g->coltab[2 * j]);
    double by = (double)(-g->rowtab[2 * i + 1]);
    double bz = (double)(g->rowtab[2 * i] * g->coltab[2 * j + 1]);
    double dd = (double)dep;
    double q0 = dd * bx - pose[9], q1 = dd * by - pose[10], q2 = dd * bz - pose[11];
    const double* R = pose;
    X[0] = R[0] * q0 + R[3] * q1 + R[6] * q2;
    X[1] = R[1] * q0 + R[4] * q1 + R[7] * q2;
    X[2] = R[2] * q0 + R[5] * q1 + R[8] * q2;
}

/* ReprojectEquirect (depthmap.cc:143-161) + the depth cv::norm of the caller:
 * returns 1 if too near (|X| < 1e-4); else (x, y) and the float distance */
static inline int reproject(const geo_t* g, const double* pose, const double* Xw, float* x, float* y, float* dist) {
    const double* R = pose; const double* t = pose + 9;
    double c0 = R[0] * Xw[0] + R[1] * Xw[1] + R[2] * Xw[2] + t[0];
    double c1 = R[3] * Xw[0] + R[4] * Xw[1] + R[5] * Xw[2] + t[1];
    double c2 = R[6] * Xw[0] + R[7] * Xw[1] + R[8] * Xw[2] + t[2];
    double r2 = c0 * c0 + c1 * c1 + c2 * c2;
    if (r2 < 1e-8) return 1;
    float xf = (float)c0, yf = (float)c1, zf = (float)c2;
    float r2f = or_clamp(fmaf(xf, xf, fmaf(yf, yf, zf * zf)), 1e-30f, 1e30f);
    float ir = or_rsqrt_pos(r2f);
    float sy = or_clamp(yf * ir, -1.0f, 1.0f);
    *y = fmaf(or_asin(sy), g->kv, g->hv);
    *x = fmaf(or_atan2(xf, zf), g->ku, g->hu);
    *dist = r2f * ir;
    return 0;
}

/* LinearInterpolation<float> (depthmap.cc:34-50) */
static inline float lin_interp(const float* img, int W, int H, float y, float x) {
    if (x < 0.0f || x >= (float)(W - 1) || y < 0.0f || y >= (float)(H - 1)) return 0.0f;
    int ix = (int)x, iy = (int)y;
    float dx = x - (float)ix, dy = y - (float)iy;
    float im00 = img[(size_t)iy * W + ix], im01 = img[(size_t)iy * W + ix + 1];
    float im10 = img[(size_t)(iy + 1) * W + ix], im11 = img[(size_t)(iy + 1) * W + ix + 1];
    float im0 = (1.0f - dx) * im00 + dx * im01;
    float im1 = (1.0f - dx) * im10 + dx * im11;
    return (1.0f - dy) * im0 + dy * im1;
}

/* DepthmapCleaner::Clean (depthmap.cc:1320-1377).  depths[0] is the map being
 * cleaned (the reference swaps its centre map to index 0 first), poses n*12. */
EXPORT int or_clean(int W, int H, int n, const float* const* depths, const double* poses,
                    float same_depth_thr, int min_consistent, float* out) {
    geo_t g; geo_init(&g, W, H);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            size_t k = (size_t)i * W + j;
            float dep = depths[0][k];
            out[k] = 0.0f;
            if ((double)dep < 1e-6) continue;
            double Xw[3];
            backproject(&g, poses, i, j, dep, Xw);
            int consistent = 1;
            for (int o = 1; o < n; ++o) {
                float x, y, dp;
                if (reproject(&g, poses + 12 * o, Xw, &x, &y, &dp)) continue;
                int iy = (int)y, ix = (int)x;
                if (!(iy >= 0 && iy < H && ix >= 0 && ix < W)) continue;
                float dr = lin_interp(depths[o], W, H, y, x);
                if (fabsf(dr - dp) < dp * same_depth_thr) consistent++;
            }
            if (consistent >= min_consistent) out[k] = dep;
        }
    geo_free(&g);
    return 0;
}

/* or_clean with the two optional per-view tests of pmp_clean_ex (no
 * counterpart in the reference): a round trip of the point map o holds at
 * the reprojection back into map 0 within sqrt(px2) pixels, and world-frame
 * normals within max_normal_deg.  planes: n maps of W*H*4 floats (camera
 * frame p = n/d), needed when max_normal_deg > 0.  Thresholds <= 0: test off. */
static int clean_round_trip(const geo_t* g, const double* pose0, const double* poseo, int i, int j, float x, float y,
                            float dr, float iw, float ih, float px2) {
    /* bearing at the continuous pixel (x, y): lambda = 2 pi (x/W - 1/2), phi = pi (1/2 - y/H), as turns */
    float ul = fmaf(x, iw, 0.5f);
    ul = ul - floorf(ul);
    float up = fmaf(y, -0.5f * ih, 0.25f);
    if (up < 0.0f) up = up + 1.0f;
    float sl, cl, sp, cp;
    or_sincos2pi(ul, &sl, &cl);
    or_sincos2pi(up, &sp, &cp);
    const double b0 = (double)(cp * sl), b1 = (double)(-sp), b2 = (double)(cp * cl);
    const double dd = (double)dr;
    const double q0 = dd * b0 - poseo[9], q1 = dd * b1 - poseo[10], q2 = dd * b2 - poseo[11];
    double X[3];
    X[0] = poseo[0] * q0 + poseo[3] * q1 + poseo[6] * q2;
    X[1] = poseo[1] * q0 + poseo[4] * q1 + poseo[7] * q2;
    X[2] = poseo[2] * q0 + poseo[5] * q1 + poseo[8] * q2;
    float x0, y0, d0;
    if (reproject(g, pose0, X, &x0, &y0, &d0)) return 0;
    const float Wf = (float)g->W;
    float dx = x0 - (float)j;
    if (dx > 0.5f * Wf) dx = dx - Wf;
    else if (dx < -0.5f * Wf) dx = dx + Wf;
    const float dy = y0 - (float)i;
    return fmaf(dx, dx, dy * dy) <= px2;
}

static int clean_normal(const geo_t* g, const double* R, const double* S, const float* plane0, const float* planeo,
                        size_t k, float x, float y, double cos2) {
    int ix = (int)(x + 0.5f), iy = (int)(y + 0.5f);
    if (ix > g->W - 1) ix = g->W - 1;
    if (iy > g->H - 1) iy = g->H - 1;
    const float* p = plane0 + 4 * k;
    const float* q = planeo + 4 * ((size_t)iy * g->W + ix);
    const double p0 = p[0], p1 = p[1], p2 = p[2], q0 = q[0], q1 = q[1], q2 = q[2];
    const double a0 = R[0] * p0 + R[3] * p1 + R[6] * p2;
    const double a1 = R[1] * p0 + R[4] * p1 + R[7] * p2;
    const double a2 = R[2] * p0 + R[5] * p1 + R[8] * p2;
    const double b0 = S[0] * q0 + S[3] * q1 + S[6] * q2;
    const double b1 = S[1] * q0 + S[4] * q1 + S[7] * q2;
    const double b2 = S[2] * q0 + S[5] * q1 + S[8] * q2;
    const double d = a0 * b0 + a1 * b1 + a2 * b2;
    if (!(d > 0.0)) return 0;
    const double na = a0 * a0 + a1 * a1 + a2 * a2, nb = b0 * b0 + b1 * b1 + b2 * b2;
    return d * d >= cos2 * (na * nb);
}

EXPORT int or_clean_ex(int W, int H, int n, const float* const* depths, const float* const* planes,
                       const double* poses, float same_depth_thr, int min_consistent, float max_reproj_px,
                       float max_normal_deg, float* out) {
    const float px2 = max_reproj_px > 0.0f ? max_reproj_px * max_reproj_px : 0.0f;
    double cos2 = 0.0;
    if (max_normal_deg > 0.0f) {
        const double cs = cos((double)max_normal_deg * (M_PI / 180.0));
        cos2 = cs * cs;
    }
    geo_t g; geo_init(&g, W, H);
    const float iw = (float)(1.0 / W), ih = (float)(1.0 / H);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            size_t k = (size_t)i * W + j;
            float dep = depths[0][k];
            out[k] = 0.0f;
            if ((double)dep < 1e-6) continue;
            double Xw[3];
            backproject(&g, poses, i, j, dep, Xw);
            int consistent = 1;
            for (int o = 1; o < n; ++o) {
                float x, y, dp;
                if (reproject(&g, poses + 12 * o, Xw, &x, &y, &dp)) continue;
                int iy = (int)y, ix = (int)x;
                if (!(iy >= 0 && iy < H && ix >= 0 && ix < W)) continue;
                float dr = lin_interp(depths[o], W, H, y, x);
                if (!(fabsf(dr - dp) < dp * same_depth_thr)) continue;
                if (px2 > 0.0f && !clean_round_trip(&g, poses, poses + 12 * o, i, j, x, y, dr, iw, ih, px2)) continue;
                if (cos2 > 0.0 && !clean_normal(&g, poses, poses + 12 * o, planes[0], planes[o], k, x, y, cos2))
                    continue;
                consistent++;
            }
            if (consistent >= min_consistent) out[k] = dep;
        }
    geo_free(&g);
    return 0;
}

/* DepthmapPruner::Prune (depthmap.cc:1535-1696), sequential raster order.
 * depths[0] is the head map, depths[1..n-1] the later maps, which are MODIFIED
 * (pixels zeroed) exactly as the reference does.  Emits points in raster
 * order into pts (3 doubles), cols (3 bytes from bgr_head), kps (i, j), up to
 * cap; *count receives the total number of kept points. */
EXPORT int or_prune(int W, int H, int n, float* const* depths, const double* poses, const uint8_t* bgr_head,
                    float same_depth_thr, float* pruned, int64_t cap, double* pts, uint8_t* cols,
                    int32_t* kps, int64_t* count) {
    geo_t g; geo_init(&g, W, H);
    const float* dm = depths[0];
    int64_t np = 0;
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            size_t k = (size_t)i * W + j;
            float dep = dm[k];
            int filt = (dep == 0.0f);
            /* :1565-1576, including the i-1>0 / j-1>0 tests */
            if (i - 1 > 0 && i - 1 < H) filt = filt || dm[k - W] == 0.0f;
            if (j - 1 > 0 && j - 1 < W) filt = filt || dm[k - 1] == 0.0f;
            if (i + 1 > 0 && i + 1 < H) filt = filt || dm[k + W] == 0.0f;
            if (j + 1 > 0 && j + 1 < W) filt = filt || dm[k + 1] == 0.0f;
            if (filt) { pruned[k] = 0.0f; continue; }
            double Xw[3];
            backproject(&g, poses, i, j, dep, Xw);
            pruned[k] = dep;
            int keep = 1;
            for (int o = 1; o < n; ++o) {
                float x, y, dp;
                if (reproject(&g, poses + 12 * o, Xw, &x, &y, &dp)) continue;
                int iu = (int)x, iv = (int)y;
                if (!(iv >= 0 && iv < H && iu >= 0 && iu < W) || (double)dp < 1e-5) continue;
                float* other = depths[o];
                float dr = other[(size_t)iv * W + iu];
                if (dr == 0.0f || fabsf(dr - dp) > same_depth_thr * dp) continue;
                if (dep < dr) {
                    other[(size_t)iv * W + iu] = 0.0f;
                } else {
                    pruned[k] = 0.0f;
                    keep = 0;
                    break;
                }
            }
            if (keep) {
                if (np < cap) {
                    pts[3 * np] = Xw[0]; pts[3 * np + 1] = Xw[1]; pts[3 * np + 2] = Xw[2];
                    cols[3 * np] = bgr_head[3 * k]; cols[3 * np + 1] = bgr_head[3 * k + 1]; cols[3 * np + 2] = bgr_head[3 * k + 2];
                    kps[2 * np] = i; kps[2 * np + 1] = j;
                }
                ++np;
            }
        }
    *count = np;
    geo_free(&g);
    return 0;
}

/* ------------------------------------------------------------------ */
/* FP64 scorer: the same warp, sampling and NCC in double, with libm    */
/* functions, exact bearings, the two-step transform of :124-130, the   */
/* single-exp bilateral weight of :149 and raw one-pass moments.        */
/* ------------------------------------------------------------------ */
static double sample_d(const uint8_t* img, int W, int H, int edge, double u, double v) {
    if (edge == OR_EDGE_ZERO) {
        if (!(u >= 0.0 && u < W - 1 && v >= 0.0 && v < H - 1)) return 0.0;
        int x0 = (int)floor(u), y0 = (int)floor(v);
        double fx = u - x0, fy = v - y0;
        const uint8_t* r0 = img + (size_t)y0 * W; const uint8_t* r1 = r0 + W;
        double top = r0[x0] + fx * (r0[x0 + 1] - r0[x0]);
        double bot = r1[x0] + fx * (r1[x0 + 1] - r1[x0]);
        return top + fy * (bot - top);
    }
    double uu = (u >= 0.0) ? (u <= (edge == OR_EDGE_WRAP ? W : W - 1) ? u : (edge == OR_EDGE_WRAP ? W : W - 1)) : 0.0;
    double vv = (v >= 0.0) ? (v <= H - 1 ? v : H - 1) : 0.0;
    int x0 = (int)floor(uu), y0 = (int)floor(vv);
    double fx = uu - x0, fy = vv - y0;
    if (x0 >= W) x0 -= W;
    int x1 = x0 + 1;
    if (x1 >= W) x1 = (edge == OR_EDGE_WRAP) ? 0 : W - 1;
    int y1 = (y0 + 1 < H) ? y0 + 1 : H - 1;
    const uint8_t* r0 = img + (size_t)y0 * W; const uint8_t* r1 = img + (size_t)y1 * W;
    double top = r0[x0] + fx * (r0[x1] - r0[x0]);
    double bot = r1[x0] + fx * (r1[x1] - r1[x0]);
    return top + fy * (bot - top);
}

/* exact (u, v) of tap t of pixel (i, j) through plane p in view v: bearings
 * from double trig, the two-step rigid transform of PatchMatchCuda.cu:117-130
 * (|b.p| clamped to >= 1e-6), libm atan2/asin (:132-144) */
static void warp_tap_d(const or_keyframe* kf, const derived_t* d, int ii, int jj, int v, double px, double py,
                       double pz, double* uo, double* vo) {
    int W = d->W, H = d->H;
    const double* R0 = kf->ref_pose; const double* t0 = kf->ref_pose + 9;
    const double* Rn = kf->view_poses + 12 * v; const double* tn = Rn + 9;
    double lon = 2.0 * M_PI * ((double)jj / W - 0.5);
    double lat = -M_PI * ((double)ii / H - 0.5);
    double bx = cos(lat) * sin(lon), by = -sin(lat), bz = cos(lat) * cos(lon);
    double s = px * bx + py * by + pz * bz;
    double as = fabs(s); if (!(as >= 1e-6)) as = 1e-6;
    if (s < 0) as = -as;
    double X[3] = { bx / as - t0[0], by / as - t0[1], bz / as - t0[2] };
    double Xw[3];
    for (int r = 0; r < 3; ++r) Xw[r] = R0[r] * X[0] + R0[3 + r] * X[1] + R0[6 + r] * X[2];
    double Xn[3];
    for (int r = 0; r < 3; ++r) Xn[r] = Rn[3 * r] * Xw[0] + Rn[3 * r + 1] * Xw[1] + Rn[3 * r + 2] * Xw[2] + tn[r];
    double rr = sqrt(Xn[0] * Xn[0] + Xn[1] * Xn[1] + Xn[2] * Xn[2]);
    if (rr < 1e-6) rr = 1e-6;
    double syv = Xn[1] / rr;
    if (syv > 1) syv = 1;
    if (syv < -1) syv = -1;
    *uo = W * (0.5 + atan2(Xn[0], Xn[2]) / (2.0 * M_PI));
    *vo = H * (0.5 + asin(syv) / M_PI);
}

/* FP64 score of plane (px, py, pz) at pixel (i, j): single-exp bilateral
 * weight (:149), raw one-pass moments (:150-169; for REF_FOLD the sums run on
 * across views, :57-62, :101), NCC with the variance floor, aggregation.
 * uv32 (nullable): sample view v's tap t at the given FP32 coordinates
 * uv32[2 (v T + t)], instead of the exact warp. */
static double score_pixel_d(const or_keyframe* kf, const or_params* p, const derived_t* d, int i, int j, double px,
                            double py, double pz, const float* uv32, double* view_scores) {
    int W = d->W, H = d->H, N = d->N;
    double cf = 1.0 / (2.0 * (double)p->sigma_color * p->sigma_color);
    double xf = 1.0 / (2.0 * (double)p->sigma_space * p->sigma_space);
    double Ic = kf->ref[(size_t)i * W + j];
    double sv[OR_MAX_VIEWS];
    const int pooled = p->agg_mode == OR_AGG_REF_FOLD;
    double sx = 0, sy_ = 0, sxx = 0, syy = 0, sxy = 0, sw = 0;
    for (int v = 0; v < N; ++v) {
        if (!pooled) { sx = 0; sy_ = 0; sxx = 0; syy = 0; sxy = 0; sw = 0; }
        for (int t = 0; t < d->T; ++t) {
            int ii = i + d->k1[t], jj = j + d->k2[t];
            double u, vv;
            if (uv32) { u = uv32[2 * (v * d->T + t)]; vv = uv32[2 * (v * d->T + t) + 1]; }
            else warp_tap_d(kf, d, ii, jj, v, px, py, pz, &u, &vv);
            double y = sample_d(kf->views[v], W, H, p->edge_mode, u, vv);
            double x = kf->ref[(size_t)ii * W + jj];
            double w = exp(-(x - Ic) * (x - Ic) * cf - (double)(d->k1[t] * d->k1[t] + d->k2[t] * d->k2[t]) * xf);
            sx += w * x; sy_ += w * y; sxx += w * x * x; syy += w * y * y; sxy += w * x * y; sw += w;
        }
        double mx = sx / sw, my = sy_ / sw;
        double vx = sxx / sw - mx * mx, vy = syy / sw - my * my;
        double c = (vx < p->var_floor || vy < p->var_floor) ? -1.0 : (sxy / sw - mx * my) / sqrt(vx * vy);
        if (c > 1) c = 1;
        if (c < -1) c = -1;
        sv[v] = c;
        if (view_scores) view_scores[v] = c;
    }
    double s;
    if (p->agg_mode == OR_AGG_MAX) {
        s = -1; for (int v = 0; v < N; ++v) if (sv[v] > s) s = sv[v];
    } else if (p->agg_mode == OR_AGG_TOPK_MEAN) {
        double tmp[OR_MAX_VIEWS]; for (int v = 0; v < N; ++v) tmp[v] = sv[v];
        s = 0;
        for (int a = 0; a < d->topk; ++a) {
            int best = a; for (int b = a + 1; b < N; ++b) if (tmp[b] > tmp[best]) best = b;
            double t = tmp[a]; tmp[a] = tmp[best]; tmp[best] = t; s += tmp[a];
        }
        s /= d->topk;
    } else {   /* REF_FOLD (pooled per-view scores), FOLD_INDEP */
        s = -1;
        for (int v = 0; v < N; ++v) {
            if (fabs(s - sv[v]) < p->fold_agree * fabs(sv[v])) s = 0.5 * (s + sv[v]);
            else if (sv[v] > s) s = sv[v];
        }
    }
    return s;
}

EXPORT int or_score_field_f64(const or_keyframe* kf, const or_params* p, const float* plane4, double* score,
                              double* view_scores) {
    derived_t d;
    if (derive(kf, p, &d)) return -1;
    int W = d.W, H = d.H, N = d.N;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            size_t k = (size_t)i * W + j;
            if (!interior(&d, i, j)) {
                score[k] = NAN;
                if (view_scores) for (int v = 0; v < N; ++v) view_scores[k * N + v] = NAN;
                continue;
            }
            score[k] = score_pixel_d(kf, p, &d, i, j, plane4[4 * k], plane4[4 * k + 1], plane4[4 * k + 2], NULL,
                                     view_scores ? view_scores + k * N : NULL);
        }
    release(&d);
    return 0;
}

/* The two halves of the score error, at n interior pixels pix[2 q] = (i, j):
 * uv32 (n N T 2 floats) gets the canonical FP32 sample coordinates of every
 * tap and view (the arithmetic of score_plane_f), uv64 (nullable, same
 * shape) the exact ones (warp_tap_d). */
EXPORT int or_warp_taps(const or_keyframe* kf, const or_params* p, const float* plane4, int64_t n, const int32_t* pix,
                        float* uv32, double* uv64) {
    derived_t d;
    if (derive(kf, p, &d)) return -1;
    int W = d.W, N = d.N, T = d.T;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t q = 0; q < n; ++q) {
        int i = pix[2 * q], j = pix[2 * q + 1];
        if (!interior(&d, i, j)) { bad = 1; continue; }
        size_t k = (size_t)i * W + j;
        float px = plane4[4 * k], py = plane4[4 * k + 1], pz = plane4[4 * k + 2];
        for (int t = 0; t < T; ++t) {
            int ii = i + d.k1[t], jj = j + d.k2[t];
            float bx, by, bz, sb[3], as;
            bearing_f(&d, ii, jj, &bx, &by, &bz);
            tap_dir_f(bx, by, bz, px, py, pz, sb, &as);
            for (int v = 0; v < N; ++v) {
                float x, y, z;
                view_dir_f(d.M[v], d.c[v], sb, as, &x, &y, &z);
                size_t o = 2 * (((size_t)q * N + v) * T + t);
                project_f(&d, x, y, z, &uv32[o], &uv32[o + 1]);
                if (uv64) warp_tap_d(kf, &d, ii, jj, v, px, py, pz, &uv64[o], &uv64[o + 1]);
            }
        }
    }
    release(&d);
    return bad ? -1 : 0;
}

/* FP64 score (score_pixel_d) at n interior pixels pix[2 q] = (i, j) with the
 * samples taken at given coordinates uv32 (n N T 2, or_warp_taps' layout):
 * the NCC arithmetic alone, with the warp's FP32 rounding factored out.
 * score (n) and view_scores (n N, nullable). */
EXPORT int or_score_at_coords_f64(const or_keyframe* kf, const or_params* p, const float* plane4, int64_t n,
                                  const int32_t* pix, const float* uv32, double* score, double* view_scores) {
    derived_t d;
    if (derive(kf, p, &d)) return -1;
    int W = d.W, N = d.N, T = d.T;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t q = 0; q < n; ++q) {
        int i = pix[2 * q], j = pix[2 * q + 1];
        if (!interior(&d, i, j)) { bad = 1; continue; }
        size_t k = (size_t)i * W + j;
        score[q] = score_pixel_d(kf, p, &d, i, j, plane4[4 * k], plane4[4 * k + 1], plane4[4 * k + 2],
                                 uv32 + 2 * (size_t)q * N * T, view_scores ? view_scores + (size_t)q * N : NULL);
    }
    release(&d);
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* diagnostics for tests                                               */
/* ------------------------------------------------------------------ */
/* which: 0 rcp 1 rsqrt 2 sqrt 3 asin 4 acos 5 atan2(a,b) 6 exp 7 log 8 sin2pi 9 cos2pi
 * 10 latitude of the direction (b, a, b) 11 longitude of the direction (a, 0, b) */
EXPORT int or_eval_math(int which, int64_t n, const float* a, const float* b, float* out) {
    for (int64_t k = 0; k < n; ++k) {
        float x = a[k], r = 0.0f, s, c;
        switch (which) {
            case 0: r = or_rcp(x); break;
            case 1: r = or_rsqrt_pos(or_clamp(x, 1e-30f, 1e30f)); break;
            case 2: r = or_sqrt(x); break;
            case 3: r = or_asin(or_clamp(x, -1.0f, 1.0f)); break;
            case 4: r = or_acos(or_clamp(x, -1.0f, 1.0f)); break;
            case 5: r = or_atan2(x, b[k]); break;
            case 6: r = or_exp(x); break;
            case 7: r = or_log(x); break;
            case 8: or_sincos2pi(x, &s, &c); r = s; break;
            case 9: or_sincos2pi(x, &s, &c); r = c; break;
            case 10: or_lonlat(b[k], x, b[k], &s, &r); break; /* latitude of (b, a, b) */
            case 11: or_lonlat(x, 0.0f, b[k], &r, &s); break;  /* longitude of (a, 0, b) */
            default: return -1;
        }
        out[k] = r;
    }
    return 0;
}

EXPORT void or_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { or_philox(ctr, key, out); }

/* the normals of refinement (pixel, iteration), n <= 3*OR_MAX_REFINE */
EXPORT int or_refine_normals(const or_keyframe* kf, const or_params* p, int i, int j, int it, int n, float* z) {
    derived_t d;
    if (derive(kf, p, &d)) return -1;
    refine_normals(&d, i, j, it, n, z);
    release(&d);
    return 0;
}

/* warped equirect coordinates (u, v) of reference pixel (ii, jj) through
 * plane p into view v (the per-tap arithmetic of score_plane_f) */
EXPORT int or_warp_point(const or_keyframe* kf, const or_params* p, int ii, int jj, int v, const float* plane,
                         float* uv) {
    derived_t d;
    if (derive(kf, p, &d)) return -1;
    if (v < 0 || v >= d.N || ii < 0 || ii >= d.H || jj < 0 || jj >= d.W) { release(&d); return -1; }
    float bx, by, bz, sb[3], as, x, y, z;
    bearing_f(&d, ii, jj, &bx, &by, &bz);
    tap_dir_f(bx, by, bz, plane[0], plane[1], plane[2], sb, &as);
    view_dir_f(d.M[v], d.c[v], sb, as, &x, &y, &z);
    project_f(&d, x, y, z, &uv[0], &uv[1]);
    release(&d);
    return 0;
}

/* view aggregation of n per-view scores under p (agg_mode, topk, fold_agree) */
EXPORT float or_aggregate(const or_params* p, int n, const float* sv) {
    derived_t d;
    d.N = n;
    int k = p->topk < 1 ? 1 : p->topk;
    if (k > n) k = n;
    d.topk = k;
    d.inv_k = (float)(1.0 / (double)k);
    return aggregate_f(&d, p, sv);
}

/* layout probe: sizeof(or_params) and the offsets of its first and last fields */
EXPORT int or_params_layout(int64_t* out) {
    out[0] = (int64_t)sizeof(or_params);
    out[1] = (int64_t)offsetof(or_params, seed);
    out[2] = (int64_t)offsetof(or_params, min_depth);
    out[3] = (int64_t)offsetof(or_params, median_rel_thr);
    return 0;
}


/* ------------------------------------------------------------------ */
/* host pre-initialisation (DepthmapEstimator::ComputePatchMatch,      */
/* depthmap.cc:690-691), restated without the OpenMP race              */
/* ------------------------------------------------------------------ */

/* PlanePosFromDepthAndNormal_equirect (depthmap.cc:130-141): n / (n.(d b)),
 * the denominator pushed away from 0 to +-1e-5 */
static inline void plane_from_depth_normal_eps(float bx, float by, float bz, float dep, float nx, float ny,
                                               float nz, float* p) {
    float qx = dep * bx, qy = dep * by, qz = dep * bz;
    float den = fmaf(nz, qz, fmaf(ny, qy, nx * qx));
    if (den < 1e-5f && den >= 0.0f) den = 1e-5f;
    else if (den > -1e-5f && den < 0.0f) den = -1e-5f;
    float r = or_rcp(den);
    p[0] = nx * r; p[1] = ny * r; p[2] = nz * r;
}

typedef struct { int pix; float plane[3]; float depth; float score; } cand_t;

/* target pixel of a world point: ReprojectEquirect (depthmap.cc:143-161) then
 * int(x + 0.5), int(y + 0.5) (:382-383, :465-466); -1 if too near, outside the
 * image or masked */
static int target_pixel(const geo_t* g, const double* pose, const double* Xw, const uint8_t* mask, float* dist) {
    float x, y;
    if (reproject(g, pose, Xw, &x, &y, dist)) return -1;
    int j = (int)(x + 0.5f), i = (int)(y + 0.5f);
    if (i < 0 || j < 0 || i >= g->H || j >= g->W) return -1;
    if (mask && !mask[(size_t)i * g->W + j]) return -1;
    return i * g->W + j;
}

/* Sequential acceptance of the candidates in index order (the loop order of
 * the reference without its OpenMP race): a candidate replaces the pixel's
 * plane/depth/score when its score is strictly higher than the pixel's
 * current score (:400, :490).  Written to the working maps and to the
 * snapshot kept for the copy-back (the reference's result maps). */
static int apply_candidates(int W, int H, const cand_t* c, int64_t n, float* plane4, float* depth, float* score,
                            float* iplane4, float* idepth, float* iscore) {
    int accepted = 0;   /* pixels changed */
    uint8_t* changed = (uint8_t*)calloc((size_t)W * H, 1);
    for (int64_t q = 0; q < n; ++q) {
        if (c[q].pix < 0) continue;
        size_t k = (size_t)c[q].pix;
        if (c[q].score > score[k]) {
            for (int a = 0; a < 3; ++a) plane4[4 * k + a] = iplane4[4 * k + a] = c[q].plane[a];
            plane4[4 * k + 3] = iplane4[4 * k + 3] = 0.0f;
            depth[k] = idepth[k] = c[q].depth;
            score[k] = iscore[k] = c[q].score;
            if (!changed[k]) ++accepted;
            changed[k] = 1;
        }
    }
    free(changed);
    return accepted;
}

static void score_candidates(const derived_t* d, const or_params* p, const or_keyframe* kf, cand_t* c, int64_t n) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < n; ++q) {
        if (c[q].pix < 0) continue;
        int i = c[q].pix / d->W, j = c[q].pix % d->W;
        refstat_t rs = ref_stats(d, kf->ref, i, j);
        c[q].score = score_plane_f(d, p, kf, &rs, i, j, c[q].plane[0], c[q].plane[1], c[q].plane[2], NULL);
    }
}

/* PreInitSparsePCL2 (depthmap.cc:361-432): each landmark (world xyz, double)
 * seeds the pixel it reprojects to with the plane facing the camera (normal
 * -bearing of that pixel, :393-395) at the landmark's distance (:388), if its
 * score beats the pixel's.  Candidates are scored with this build's scorer
 * (params' aggregation and edge mode) and only at scored pixels
 * (in_score_grid); the reference scores them with its CPU scorer
 * (ComputePlaneScore, :1104-1127).  mask: nullable.  out_scores (nullable,
 * n): candidate scores, NaN where no candidate.  Returns the number of pixels
 * changed. */
EXPORT int or_preinit_landmarks(const or_keyframe* kf, const or_params* p, float* plane4, float* depth, float* score,
                                float* iplane4, float* idepth, float* iscore, const uint8_t* mask, int64_t n,
                                const double* xyz, float* out_scores) {
    derived_t d;
    if (derive(kf, p, &d)) return -1;
    geo_t g;
    geo_init(&g, d.W, d.H);
    cand_t* c = (cand_t*)malloc(sizeof(cand_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t q = 0; q < n; ++q) {
        float dist;
        c[q].pix = target_pixel(&g, kf->ref_pose, xyz + 3 * q, mask, &dist);
        if (c[q].pix >= 0 && !in_score_grid(&d, p, c[q].pix / d.W, c[q].pix % d.W)) c[q].pix = -1;
        if (c[q].pix < 0) continue;
        float bx, by, bz;
        bearing_f(&d, c[q].pix / d.W, c[q].pix % d.W, &bx, &by, &bz);
        plane_from_depth_normal_eps(bx, by, bz, dist, -bx, -by, -bz, c[q].plane);
        c[q].depth = dist;
    }
    score_candidates(&d, p, kf, c, n);
    if (out_scores)
        for (int64_t q = 0; q < n; ++q) out_scores[q] = c[q].pix >= 0 ? c[q].score : or_u2f(NAN_BITS);
    int acc = apply_candidates(d.W, d.H, c, n, plane4, depth, score, iplane4, idepth, iscore);
    free(c);
    geo_free(&g);
    release(&d);
    return acc;
}

/* PosPatchPartOnPlane (depthmap.cc:163-177): the point of plane p on the
 * bearing of pixel (i, j) of a camera at pose, in the world frame.  scale =
 * p.b in float (the reference rounds its double dot to float), pushed away
 * from 0 to +-1e-6; the camera-frame point b/scale is formed in float, the
 * rigid transform R^T (X - t) in double. */
static inline void pos_on_plane(const geo_t* g, const double* pose, int i, int j, const float* pl, double* Xw) {
    float bx = g->rowtab[2 * i] * g->coltab[2 * j];
    float by = -g->rowtab[2 * i + 1];
    float bz = g->rowtab[2 * i] * g->coltab[2 * j + 1];
    float s = fmaf(pl[2], bz, fmaf(pl[1], by, pl[0] * bx));
    if (s < 1e-6f && s >= 0.0f) s = 1e-6f;
    else if (s > -1e-6f && s < 0.0f) s = -1e-6f;
    float r = or_rcp(s);
    double q0 = (double)(bx * r) - pose[9], q1 = (double)(by * r) - pose[10], q2 = (double)(bz * r) - pose[11];
    const double* R = pose;
    Xw[0] = R[0] * q0 + R[3] * q1 + R[6] * q2;
    Xw[1] = R[1] * q0 + R[4] * q1 + R[7] * q2;
    Xw[2] = R[2] * q0 + R[5] * q1 + R[8] * q2;
}

/* PreInitResultWarp (depthmap.cc:434-516): every pixel (r, c) of the
 * memorised previous result (MemorizeResult, :808-829: the final plane and
 * depth maps and the pose of the previous reference) with depth >= 1e-4 and
 * the CURRENT mask set at (r, c) (:445) is moved along its plane into the
 * current reference camera.  The candidate at the pixel it lands on has the
 * point's distance as depth and the previous plane's direction, normalised
 * and not rotated into the current frame (:479-481), as its normal.
 * Candidates in raster order of (r, c); same scoring and acceptance as
 * or_preinit_landmarks.  prev maps have the current keyframe's size. */
EXPORT int or_preinit_warp(const or_keyframe* kf, const or_params* p, float* plane4, float* depth, float* score,
                           float* iplane4, float* idepth, float* iscore, const uint8_t* mask, const double* prev_pose,
                           const float* prev_plane4, const float* prev_depth, float* out_scores) {
    derived_t d;
    if (derive(kf, p, &d)) return -1;
    geo_t g;
    geo_init(&g, d.W, d.H);
    int64_t n = (int64_t)d.W * d.H;
    cand_t* c = (cand_t*)malloc(sizeof(cand_t) * (size_t)n);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < n; ++q) {
        int r = (int)(q / d.W), cc = (int)(q % d.W);
        c[q].pix = -1;
        if (mask && !mask[q]) continue;
        if (!(prev_depth[q] >= 1e-4f)) continue;
        const float* pp = prev_plane4 + 4 * q;
        double Xw[3];
        pos_on_plane(&g, prev_pose, r, cc, pp, Xw);
        float dist;
        int pix = target_pixel(&g, kf->ref_pose, Xw, mask, &dist);
        if (pix < 0 || !in_score_grid(&d, p, pix / d.W, pix % d.W)) continue;
        float nr2 = or_clamp(fmaf(pp[2], pp[2], fmaf(pp[1], pp[1], pp[0] * pp[0])), 1e-30f, 1e30f);
        float inr = or_rsqrt_pos(nr2);
        float bx, by, bz;
        bearing_f(&d, pix / d.W, pix % d.W, &bx, &by, &bz);
        plane_from_depth_normal_eps(bx, by, bz, dist, pp[0] * inr, pp[1] * inr, pp[2] * inr, c[q].plane);
        c[q].depth = dist;
        c[q].pix = pix;
    }
    score_candidates(&d, p, kf, c, n);
    if (out_scores)
        for (int64_t q = 0; q < n; ++q) out_scores[q] = c[q].pix >= 0 ? c[q].score : or_u2f(NAN_BITS);
    int acc = apply_candidates(d.W, d.H, c, n, plane4, depth, score, iplane4, idepth, iscore);
    free(c);
    geo_free(&g);
    release(&d);
    return acc;
}

/* ------------------------------------------------------------------ */
/* thumbnails: the frame's grey / mask / colour thumbnails             */
/* (tracking_module.cc:82-93, data/frame.cc:53-54, :136-138), i.e.     */
/* cv::cvtColor(BGR2GRAY) then cv::resize(INTER_LINEAR) on 8-bit data. */
/* OpenCV 4.1.0 is absent here; both follow its published 8-bit        */
/* fixed-point formulas:                                               */
/*   grey  Y = (1868 B + 9617 G + 4899 R + 2^13) >> 14                 */
/*   resize: source coordinate f = (d + 0.5) * (src/dst) - 0.5 (float), */
/*   s = floor(f), f -= s, clamped to the first / last sample (f = 0);  */
/*   weights round(1 - f) * 2048 and round(f * 2048) (shorts, rounded  */
/*   independently); rows then columns in int, (v + 2^21) >> 22.       */
/* ------------------------------------------------------------------ */

/* color order: 0 grey, 1 RGB, 2 BGR (camera::color_order_t) */
EXPORT int or_to_gray(const uint8_t* src, int W, int H, int64_t pitch, int channels, int order, uint8_t* dst) {
    if (channels != 1 && channels != 3 && channels != 4) return -1;
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            const uint8_t* p = src + (size_t)i * (size_t)pitch + (size_t)j * channels;
            if (channels == 1 || order == 0) { dst[(size_t)i * W + j] = p[0]; continue; }
            int r = order == 1 ? p[0] : p[2], g = p[1], b = order == 1 ? p[2] : p[0];
            dst[(size_t)i * W + j] = (uint8_t)((b * 1868 + g * 9617 + r * 4899 + (1 << 13)) >> 14);
        }
    return 0;
}

/* per destination index: source index s0, s1 and weights a0, a1 (sum ~ 2048) */
EXPORT void or_resize_taps(int src, int dst, int32_t* s0, int32_t* s1, int16_t* a0, int16_t* a1) {
    const double scale = (double)src / (double)dst;
    for (int d = 0; d < dst; ++d) {
        float f = (float)((d + 0.5) * scale - 0.5);
        int s = (int)floorf(f);
        f -= (float)s;
        if (s < 0) { f = 0.0f; s = 0; }
        if (s >= src - 1) { f = 0.0f; s = src - 1; }
        s0[d] = s;
        s1[d] = s + 1 < src ? s + 1 : src - 1;
        a0[d] = (int16_t)lrintf((1.0f - f) * 2048.0f);
        a1[d] = (int16_t)lrintf(f * 2048.0f);
    }
}

EXPORT int or_resize_linear(const uint8_t* src, int W, int H, int64_t pitch, int channels, uint8_t* dst, int dW,
                            int dH) {
    if (W < 1 || H < 1 || dW < 1 || dH < 1 || channels < 1 || channels > 4) return -1;
    int32_t* xs0 = (int32_t*)malloc(sizeof(int32_t) * dW); int32_t* xs1 = (int32_t*)malloc(sizeof(int32_t) * dW);
    int16_t* xa0 = (int16_t*)malloc(sizeof(int16_t) * dW); int16_t* xa1 = (int16_t*)malloc(sizeof(int16_t) * dW);
    int32_t* ys0 = (int32_t*)malloc(sizeof(int32_t) * dH); int32_t* ys1 = (int32_t*)malloc(sizeof(int32_t) * dH);
    int16_t* ya0 = (int16_t*)malloc(sizeof(int16_t) * dH); int16_t* ya1 = (int16_t*)malloc(sizeof(int16_t) * dH);
    or_resize_taps(W, dW, xs0, xs1, xa0, xa1);
    or_resize_taps(H, dH, ys0, ys1, ya0, ya1);
    for (int y = 0; y < dH; ++y) {
        const uint8_t* r0 = src + (size_t)ys0[y] * (size_t)pitch;
        const uint8_t* r1 = src + (size_t)ys1[y] * (size_t)pitch;
        for (int x = 0; x < dW; ++x)
            for (int c = 0; c < channels; ++c) {
                int h0 = r0[xs0[x] * channels + c] * xa0[x] + r0[xs1[x] * channels + c] * xa1[x];
                int h1 = r1[xs0[x] * channels + c] * xa0[x] + r1[xs1[x] * channels + c] * xa1[x];
                int v = (h0 * ya0[y] + h1 * ya1[y] + (1 << 21)) >> 22;
                dst[((size_t)y * dW + x) * channels + c] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
            }
    }
    free(xs0); free(xs1); free(xa0); free(xa1); free(ys0); free(ys1); free(ya0); free(ya1);
    return 0;
}
