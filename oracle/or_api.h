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

#endif
