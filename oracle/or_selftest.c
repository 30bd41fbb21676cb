/*
 * or_selftest.c -- drives every oracle entry point on small synthetic inputs,
 * for the AddressSanitizer / UndefinedBehaviorSanitizer build of the oracle
 * (make -C oracle sanitize; run by tests/test_sanitizers.py).
 *
 * TEST INFRASTRUCTURE.  The inputs are a procedural texture seen from a few
 * translated and rotated cameras, with every aggregation, edge and coverage
 * mode, patch shapes up to the largest, masks, pre-init, the Cleaner and the
 * Pruner; the checks are return codes and plain invariants (finite depths,
 * scores in [-1, 1] or NaN where never scored).  A sanitizer report or a
 * failed check exits non-zero.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "or_api.h"

#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            fprintf(stderr, "selftest: check failed at line %d: %s\n", __LINE__, #c); \
            exit(1);                                                         \
        }                                                                    \
    } while (0)

static uint32_t hash2(uint32_t x, uint32_t y) {
    uint32_t h = x * 0x8DA6B343u ^ y * 0xD8163841u;
    h ^= h >> 13; h *= 0x85EBCA6Bu; h ^= h >> 16;
    return h;
}

/* value noise on a 8-pixel lattice plus fine noise: a textured grey image */
static void texture(uint8_t* img, int W, int H, uint32_t seed) {
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            int ci = i / 8, cj = j / 8;
            double fi = (i % 8) / 8.0, fj = (j % 8) / 8.0;
            double a = hash2(cj + seed, ci) & 255, b = hash2(cj + 1 + seed, ci) & 255;
            double c = hash2(cj + seed, ci + 1) & 255, d = hash2(cj + 1 + seed, ci + 1) & 255;
            double v = (a * (1 - fj) + b * fj) * (1 - fi) + (c * (1 - fj) + d * fj) * fi;
            v = 0.8 * v + 0.2 * (hash2(j, i + 7 * seed) & 255);
            img[(size_t)i * W + j] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
        }
}

/* world->camera pose: yaw about y, camera centre c */
static void pose(double* P, double yaw, double cx, double cy, double cz) {
    double cs = cos(yaw), sn = sin(yaw);
    double R[9] = { cs, 0, -sn, 0, 1, 0, sn, 0, cs };
    memcpy(P, R, sizeof(R));
    for (int r = 0; r < 3; ++r) P[9 + r] = -(R[3 * r] * cx + R[3 * r + 1] * cy + R[3 * r + 2] * cz);
}

static or_params defaults(void) {
    or_params p;
    memset(&p, 0, sizeof(p));
    p.patch_half = 5; p.patch_stride = 2; p.iterations = 1; p.n_refine = 6;
    p.agg_mode = OR_AGG_REF_FOLD; p.topk = 2; p.edge_mode = OR_EDGE_WRAP; p.coverage_mode = OR_COVER_REF_GRID;
    p.init_mode = OR_INIT_INVDEPTH_HEMI; p.postprocess = 1; p.seed = 12345;
    p.min_depth = 1e-5f; p.max_depth = 20.0f; p.init_inv_depth_min = 1.0f / 20.0f; p.init_inv_depth_max = 1.0f / 0.3f;
    p.min_score = 0.1f; p.sigma_color = 50.0f; p.sigma_space = 5.0f; p.var_floor = 0.1f; p.fold_agree = 0.25f;
    p.depth_penalty = 1250.0f; p.plane_penalty = 8.0f;
    p.refine_depth_range = 0.02f; p.refine_normal_range = 0.5f; p.refine_depth_decay = 0.3f;
    p.refine_normal_decay = 0.8f; p.min_patch_sd = 0.0f; p.median_rel_thr = 0.05;
    return p;
}

static void check_maps(int P, const float* depth, const float* score) {
    for (int k = 0; k < P; ++k) {
        CHECK(isfinite(depth[k]));
        CHECK(isnan(score[k]) || (score[k] >= -1.0f && score[k] <= 1.0f));
    }
}

enum { W = 96, H = 48, NV = 4 };

int main(void) {
    const int P = W * H;
    uint8_t* imgs[NV + 1];
    for (int v = 0; v <= NV; ++v) {
        imgs[v] = (uint8_t*)malloc(P);
        texture(imgs[v], W, H, 3);                 /* the same scene texture in every view */
    }
    double ref_pose[12], view_poses[12 * NV];
    pose(ref_pose, 0.1, 0.0, 0.0, 0.0);
    for (int v = 0; v < NV; ++v) pose(view_poses + 12 * v, 0.1 + 0.05 * v, 0.1 * (v + 1), 0.02 * v, -0.05 * v);
    or_keyframe kf = { W, H, NV, 0, imgs[0], (const uint8_t* const*)(imgs + 1), ref_pose, view_poses };
    uint8_t* mask = (uint8_t*)malloc(P);
    for (int k = 0; k < P; ++k) mask[k] = (k / W) < (int)(0.85 * H);

    float* plane4 = (float*)malloc(sizeof(float) * 4 * P);
    float* depth = (float*)malloc(sizeof(float) * P);
    float* score = (float*)malloc(sizeof(float) * P);
    float* ip4 = (float*)malloc(sizeof(float) * 4 * P);
    float* idp = (float*)malloc(sizeof(float) * P);
    float* isc = (float*)malloc(sizeof(float) * P);
    float* vs = (float*)malloc(sizeof(float) * P * NV);
    double* s64 = (double*)malloc(sizeof(double) * P);
    double* v64 = (double*)malloc(sizeof(double) * P * NV);
    float* od = (float*)malloc(sizeof(float) * P);
    float* op3 = (float*)malloc(sizeof(float) * 3 * P);
    float* osc = (float*)malloc(sizeof(float) * P);
    int32_t* ong = (int32_t*)malloc(sizeof(int32_t) * P);

    /* every aggregation x edge x coverage, two patch shapes, the largest patch */
    const int shapes[3][2] = { {5, 2}, {3, 1}, {10, 2} };
    for (int sh = 0; sh < 3; ++sh)
        for (int agg = 0; agg < 4; ++agg)
            for (int edge = 0; edge < 3; ++edge)
                for (int cov = 0; cov < 2; ++cov) {
                    if (sh == 2 && (agg != OR_AGG_TOPK_MEAN || cov != OR_COVER_FULL)) continue;
                    or_params p = defaults();
                    p.patch_half = shapes[sh][0]; p.patch_stride = shapes[sh][1];
                    p.agg_mode = agg; p.edge_mode = edge; p.coverage_mode = cov;
                    p.init_mode = (agg + edge) & 1;
                    p.min_patch_sd = (agg == 1) ? 4.0f : 0.0f;
                    CHECK(or_init(&kf, &p, plane4, depth, score) == 0);
                    check_maps(P, depth, score);
                    memcpy(ip4, plane4, sizeof(float) * 4 * P); memcpy(idp, depth, sizeof(float) * P);
                    memcpy(isc, score, sizeof(float) * P);
                    CHECK(or_score_field(&kf, &p, plane4, score, vs) == 0);
                    CHECK(or_score_field_f64(&kf, &p, plane4, s64, v64) == 0);
                    int64_t ev = 0;
                    CHECK(or_propagate(&kf, &p, plane4, depth, score, 0, 1, &ev) == 0);
                    CHECK(ev > 0);
                    check_maps(P, depth, score);
                    CHECK(or_finalize(&kf, &p, mask, ip4, idp, isc, plane4, depth, score, od, op3, osc, ong) == 0);
                    check_maps(P, od, osc);
                }

    /* the error split at a few pixels, one band of rows, the full estimator */
    or_params p = defaults();
    CHECK(or_init(&kf, &p, plane4, depth, score) == 0);
    {
        int32_t pix[6] = { 10, 20, 24, 48, 36, 90 };
        int T = 36;
        float* uv32 = (float*)malloc(sizeof(float) * 3 * NV * T * 2);
        double* uv64 = (double*)malloc(sizeof(double) * 3 * NV * T * 2);
        double s3[3], v3[3 * NV];
        CHECK(or_warp_taps(&kf, &p, plane4, 3, pix, uv32, uv64) == 0);
        CHECK(or_score_at_coords_f64(&kf, &p, plane4, 3, pix, uv32, s3, v3) == 0);
        int32_t bad[2] = { 0, 0 };   /* not interior: rejected */
        CHECK(or_warp_taps(&kf, &p, plane4, 1, bad, uv32, uv64) != 0);
        free(uv32); free(uv64);
    }
    CHECK(or_propagate_rows(&kf, &p, plane4, depth, score, 0, 0, 10, 30, NULL) == 0);
    CHECK(or_densify(&kf, &p, mask, od, op3, osc, ong, plane4, depth, score) == 0);
    check_maps(P, od, osc);

    /* pre-init from landmarks on a wall and from the previous result */
    {
        enum { NL = 300 };
        double xyz[3 * NL];
        for (int l = 0; l < NL; ++l) {
            double a = 6.283185307 * l / NL;
            xyz[3 * l] = 3.0 * sin(a); xyz[3 * l + 1] = 0.3 * cos(7 * a); xyz[3 * l + 2] = 3.0 * cos(a);
        }
        float* sc_out = (float*)malloc(sizeof(float) * NL);
        CHECK(or_init(&kf, &p, plane4, depth, score) == 0);
        memcpy(ip4, plane4, sizeof(float) * 4 * P); memcpy(idp, depth, sizeof(float) * P);
        memcpy(isc, score, sizeof(float) * P);
        CHECK(or_preinit_landmarks(&kf, &p, plane4, depth, score, ip4, idp, isc, mask, NL, xyz, sc_out) > 0);
        float* prev4 = (float*)malloc(sizeof(float) * 4 * P);
        float* prevd = (float*)malloc(sizeof(float) * P);
        memcpy(prev4, plane4, sizeof(float) * 4 * P); memcpy(prevd, depth, sizeof(float) * P);
        float* wsc = (float*)malloc(sizeof(float) * P);
        CHECK(or_preinit_warp(&kf, &p, plane4, depth, score, ip4, idp, isc, mask, view_poses, prev4, prevd, wsc) >= 0);
        check_maps(P, depth, score);
        free(sc_out); free(prev4); free(prevd); free(wsc);
    }

    /* Cleaner (both forms) and Pruner over five maps of one wall */
    {
        enum { NM = 5 };
        float* maps[NM];
        float* planes[NM];
        double poses[12 * NM];
        for (int m = 0; m < NM; ++m) {
            pose(poses + 12 * m, 0.02 * m, 0.05 * m, 0.0, 0.0);
            maps[m] = (float*)malloc(sizeof(float) * P);
            planes[m] = (float*)malloc(sizeof(float) * 4 * P);
            for (int k = 0; k < P; ++k) {
                maps[m][k] = (k % 17 == 0) ? 0.0f : 2.0f + 0.001f * (float)(k % 5);
                planes[m][4 * k] = 0.0f; planes[m][4 * k + 1] = 0.0f; planes[m][4 * k + 2] = 0.5f;
                planes[m][4 * k + 3] = 0.0f;
            }
        }
        float* out = (float*)malloc(sizeof(float) * P);
        CHECK(or_clean(W, H, NM, (const float* const*)maps, poses, 0.03f, 3, out) == 0);
        CHECK(or_clean_ex(W, H, NM, (const float* const*)maps, (const float* const*)planes, poses, 0.03f, 3, 1.0f,
                          10.0f, out) == 0);
        uint8_t* bgr = (uint8_t*)malloc(3 * P);
        memset(bgr, 90, 3 * P);
        float* pruned = (float*)malloc(sizeof(float) * P);
        double* pts = (double*)malloc(sizeof(double) * 3 * P);
        uint8_t* cols = (uint8_t*)malloc(3 * P);
        int32_t* kps = (int32_t*)malloc(sizeof(int32_t) * 2 * P);
        int64_t count = 0;
        CHECK(or_prune(W, H, NM, maps, poses, bgr, 0.03f, pruned, P, pts, cols, kps, &count) == 0);
        CHECK(count >= 0 && count <= P);
        for (int m = 0; m < NM; ++m) { free(maps[m]); free(planes[m]); }
        free(out); free(bgr); free(pruned); free(pts); free(cols); free(kps);
    }

    /* elementary functions, random streams, aggregation, thumbnails */
    {
        enum { NX = 4096 };
        float a[NX], b[NX], out[NX];
        for (int k = 0; k < NX; ++k) { a[k] = (float)(k - NX / 2) / 512.0f; b[k] = (float)(NX - k) / 700.0f; }
        for (int which = 0; which <= 10; ++which) {
            float aa[NX];
            for (int k = 0; k < NX; ++k) aa[k] = (which == 7) ? (float)(k + 1) / NX : a[k];
            CHECK(or_eval_math(which, NX, aa, b, out) == 0);
        }
        uint32_t ctr[4] = { 1, 2, 3, 4 }, key[2] = { 5, 6 }, r[4];
        or_philox4x32_10(ctr, key, r);
        float z[18];
        CHECK(or_refine_normals(&kf, &p, 5, 7, 1, 18, z) == 0);
        float sv[3] = { 0.2f, 0.9f, -0.4f };
        for (int agg = 0; agg < 4; ++agg) {
            or_params q = defaults();
            q.agg_mode = agg;
            float s = or_aggregate(&q, 3, sv);
            CHECK(s >= -1.0f && s <= 1.0f);
        }
        float uv[2], pl[3] = { 0.0f, 0.0f, 0.5f };
        CHECK(or_warp_point(&kf, &p, 20, 30, 1, pl, uv) == 0);
        int64_t lay[4];
        CHECK(or_params_layout(lay) == 0 && lay[0] == (int64_t)sizeof(or_params));
        uint8_t* rgb = (uint8_t*)malloc(3 * P);
        for (int k = 0; k < 3 * P; ++k) rgb[k] = (uint8_t)hash2(k, 1);
        uint8_t* grey = (uint8_t*)malloc(P);
        CHECK(or_to_gray(rgb, W, H, 3 * W, 3, 2, grey) == 0);
        uint8_t* small = (uint8_t*)malloc(3 * 37 * 19);
        CHECK(or_resize_linear(rgb, W, H, 3 * W, 3, small, 37, 19) == 0);
        uint8_t* big = (uint8_t*)malloc(3 * 150 * 70);
        CHECK(or_resize_linear(rgb, W, H, 3 * W, 3, big, 150, 70) == 0);
        float rt[2 * H], ct[2 * W], M[9 * NV], c[3 * NV];
        CHECK(or_tables(W, H, rt, ct) == 0);
        CHECK(or_derive_poses(ref_pose, NV, view_poses, M, c) == 0);
        free(rgb); free(grey); free(small); free(big);
    }

    /* argument errors come back as -1 */
    {
        or_params q = defaults();
        q.patch_half = 2;
        CHECK(or_init(&kf, &q, plane4, depth, score) != 0);
        q = defaults();
        q.n_refine = 99;
        CHECK(or_propagate(&kf, &q, plane4, depth, score, 0, 1, NULL) != 0);
    }

    for (int v = 0; v <= NV; ++v) free(imgs[v]);
    free(mask); free(plane4); free(depth); free(score); free(ip4); free(idp); free(isc); free(vs);
    free(s64); free(v64); free(od); free(op3); free(osc); free(ong);
    printf("or_selftest: ok\n");
    return 0;
}
