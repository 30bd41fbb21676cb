/*
 * or_api.h -- C interface of the CPU oracle (liboracle.so).
 *
 * TEST INFRASTRUCTURE.  The oracle is a plain-C restatement of the
 * reference's densification path (src/openvslam/PatchMatchCuda.cu and
 * src/openvslam/depthmap.cc) written to this project's canonical FP32
 * semantics, plus an FP64 scorer.  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may load it.
 *
 * Struct layouts match include/pmp/pmp.h field for field, so one ctypes
 * description serves both libraries.
 */
#ifndef OR_API_H
#define OR_API_H
#include <stdint.h>

#define OR_MAX_VIEWS 16
#define OR_MAX_TAPS 121
#define OR_MAX_REFINE 8

/* REF_FOLD: the device fold of PatchMatchCuda.cu:161-171 over moments pooled
 * across views (:57-62 zero them once, outside the view loop :101);
 * FOLD_INDEP: the same fold over independent per-view NCCs */
enum { OR_AGG_REF_FOLD = 0, OR_AGG_MAX = 1, OR_AGG_TOPK_MEAN = 2, OR_AGG_FOLD_INDEP = 3 };
enum { OR_EDGE_WRAP = 0, OR_EDGE_CLAMP = 1, OR_EDGE_ZERO = 2 };
enum { OR_COVER_REF_GRID = 0, OR_COVER_FULL = 1 };
enum { OR_INIT_REF_UNIFORM = 0, OR_INIT_INVDEPTH_HEMI = 1 };

typedef struct {
    int32_t patch_half, patch_stride, iterations, n_refine;
    int32_t agg_mode, topk, edge_mode, coverage_mode;
    int32_t init_mode, postprocess;
    uint64_t seed;
    float min_depth, max_depth;
    float init_inv_depth_min, init_inv_depth_max;
    float min_score, sigma_color, sigma_space, var_floor, fold_agree;
    float depth_penalty, plane_penalty;
    float refine_depth_range, refine_normal_range, refine_depth_decay, refine_normal_decay;
    float min_patch_sd;
    double median_rel_thr;   /* compared in double, as the literal 0.05 of depthmap.cc:1190 */
} or_params;

typedef struct {
    int32_t W, H, N, pad_;
    const uint8_t* ref;            /* H*W grey, row pitch W */
    const uint8_t* const* views;   /* N pointers, H*W each */
    const double* ref_pose;        /* 12: R (row-major, world->camera), t */
    const double* view_poses;      /* N*12 */
} or_keyframe;

/* entry points (or_core.c); 0 on success, -1 on a bad argument */
int or_tables(int W, int H, float* rowtab, float* coltab);
int or_derive_poses(const double* ref_pose, int N, const double* view_poses, float* M, float* c);
int or_init(const or_keyframe* kf, const or_params* p, float* plane4, float* depth, float* score);
int or_score_field(const or_keyframe* kf, const or_params* p, const float* plane4, float* score, float* view_scores);
int or_score_field_f64(const or_keyframe* kf, const or_params* p, const float* plane4, double* score,
                       double* view_scores);
int or_warp_taps(const or_keyframe* kf, const or_params* p, const float* plane4, int64_t n, const int32_t* pix,
                 float* uv32, double* uv64);
int or_score_at_coords_f64(const or_keyframe* kf, const or_params* p, const float* plane4, int64_t n,
                           const int32_t* pix, const float* uv32, double* score, double* view_scores);
int or_propagate(const or_keyframe* kf, const or_params* p, float* plane4, float* depth, float* score, int it_begin,
                 int it_end, int64_t* evals_out);
int or_propagate_rows(const or_keyframe* kf, const or_params* p, float* plane4, float* depth, float* score, int it,
                      int colour, int row_begin, int row_end, int64_t* evals_out);
int or_finalize(const or_keyframe* kf, const or_params* p, const uint8_t* mask, const float* init_plane4,
                const float* init_depth, const float* init_score, const float* fin_plane4, const float* fin_depth,
                const float* fin_score, float* out_depth, float* out_plane3, float* out_score, int32_t* out_nghbr);
int or_densify(const or_keyframe* kf, const or_params* p, const uint8_t* mask, float* out_depth, float* out_plane3,
               float* out_score, int32_t* out_nghbr, float* raw_plane4, float* raw_depth, float* raw_score);
int or_clean(int W, int H, int n, const float* const* depths, const double* poses, float same_depth_thr,
             int min_consistent, float* out);
int or_clean_ex(int W, int H, int n, const float* const* depths, const float* const* planes, const double* poses,
                float same_depth_thr, int min_consistent, float max_reproj_px, float max_normal_deg, float* out);
int or_prune(int W, int H, int n, float* const* depths, const double* poses, const uint8_t* bgr_head,
             float same_depth_thr, float* pruned, int64_t cap, double* pts, uint8_t* cols, int32_t* kps,
             int64_t* count);
int or_eval_math(int which, int64_t n, const float* a, const float* b, float* out);
void or_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out);
int or_refine_normals(const or_keyframe* kf, const or_params* p, int i, int j, int it, int n, float* z);
int or_warp_point(const or_keyframe* kf, const or_params* p, int ii, int jj, int v, const float* plane, float* uv);
float or_aggregate(const or_params* p, int n, const float* sv);
int or_params_layout(int64_t* out);
int or_preinit_landmarks(const or_keyframe* kf, const or_params* p, float* plane4, float* depth, float* score,
                         float* iplane4, float* idepth, float* iscore, const uint8_t* mask, int64_t n,
                         const double* xyz, float* out_scores);
int or_preinit_warp(const or_keyframe* kf, const or_params* p, float* plane4, float* depth, float* score,
                    float* iplane4, float* idepth, float* iscore, const uint8_t* mask, const double* prev_pose,
                    const float* prev_plane4, const float* prev_depth, float* out_scores);
int or_to_gray(const uint8_t* src, int W, int H, int64_t pitch, int channels, int order, uint8_t* dst);
void or_resize_taps(int src, int dst, int32_t* s0, int32_t* s1, int16_t* a0, int16_t* a1);
int or_resize_linear(const uint8_t* src, int W, int H, int64_t pitch, int channels, uint8_t* dst, int dW, int dH);

#endif
