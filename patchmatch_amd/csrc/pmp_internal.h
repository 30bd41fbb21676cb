// pmp_internal.h -- context, device buffers and kernel argument blocks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <unordered_map>
#include <vector>

#include "pmp/pmp.h"

namespace pmp {

constexpr int kMaxViews = PMP_MAX_VIEWS;
constexpr int kMaxTaps = 121;   // patch_half 5 at stride 1
constexpr int kMaxRefine = 8;

// Kernel argument block of one keyframe: everything uniform across the grid.
// Passed by value (kernarg segment, read with scalar loads).
struct KfArgs {
    int W, H, N, half, stride, T;
    int agg_mode, topk, edge_mode, coverage, init_mode, n_refine;
    const uint8_t* ref;                   // H*W
    const uint32_t* quads;                // N*H*W packed 2x2 neighbourhoods, view v at v*H*W
    const float2* rowtab;                 // H: (cos lat, sin lat)
    const float2* coltab;                 // W: (sin lon, cos lon)
    const float* Wc;                      // 256: exp(-a^2 / (2 sigma_color^2))
    float Ws[kMaxTaps];                   // exp(-|k|^2 / (2 sigma_space^2))
    float M[kMaxViews][9];                // R_n R_ref^T
    float c[kMaxViews][3];                // t_n - M t_ref
    float ku, hu, kv, hv;                 // u = lon*ku + hu, v = asin*kv + hv
    float umax, vmax;                     // sample clamps: W (wrap) or W-1, and H-1
    uint32_t quads_bytes;                 // size of the packed-view buffer
    float var_floor, fold_agree, inv_k;
    float depth_penalty, plane_penalty;
    float dr[kMaxRefine], nr[kMaxRefine];
    float min_depth, max_depth, inv_lo, inv_hi;
    uint32_t key0, key1;
};

// Maps of one keyframe workspace (device).
struct Maps {
    float4* plane = nullptr;
    float* depth = nullptr;
    float* score = nullptr;
};

template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t n = 0;
};

struct Image {
    int W = 0, H = 0;
    uint8_t* gray = nullptr;          // H*W
    size_t bytes = 0;
    uint64_t generation = 0;          // changes whenever the id is re-uploaded
};

struct PrunerView {
    int W = 0, H = 0;
    float* depth = nullptr;   // device, modified by pruning as in the reference
    uint8_t* bgr = nullptr;   // device, H*W*3
    double pose[12];
};

}  // namespace pmp

struct pmp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;

    std::unordered_map<uint64_t, pmp::Image> images;

    // current keyframe
    bool views_set = false;
    uint64_t ref_id = 0;
    std::vector<uint64_t> view_ids;
    double ref_pose[12];
    int64_t mask_id = -1;             // pooled image used as mask (-1: none)
    std::vector<double> view_poses;   // N*12
    int W = 0, H = 0;

    // workspace (grown on demand, never freed per call)
    size_t ws_pixels = 0;
    pmp::Maps work, init;
    float4* out_plane = nullptr;
    float* out_depth = nullptr;
    float* out_score = nullptr;
    int32_t* out_nghbr = nullptr;
    uint32_t* views_q = nullptr;      // packed neighbours of the current keyframe, contiguous
    size_t views_q_cap = 0;           // in words
    std::vector<uint64_t> packed_key; // (id, generation) pairs + edge flag of views_q
    uint64_t next_generation = 1;
    float* tmp0 = nullptr;
    float* tmp1 = nullptr;
    uint8_t* mask = nullptr;
    float2* rowtab = nullptr;
    float2* coltab = nullptr;
    int tab_W = 0, tab_H = 0;
    float* wc = nullptr;
    float wc_sigma = -1.0f;
    unsigned long long* counters = nullptr;   // [0] evals, [1] candidates
    bool tables_ok = false;

    // cleaner
    float* clean_maps = nullptr;
    size_t clean_cap = 0;
    double* clean_poses = nullptr;

    // pruner
    std::vector<pmp::PrunerView> pruner;
    int32_t* prune_z = nullptr;        // per target pixel: first zeroing head pixel
    uint8_t* prune_alive = nullptr;
    float* prune_out = nullptr;
    double* prune_pts = nullptr;
    uint8_t* prune_cols = nullptr;
    int32_t* prune_kps = nullptr;
    int32_t* prune_blk = nullptr;
    double* prune_poses = nullptr;
    size_t prune_cap = 0;

    // stage timing (HIP events on `stream`)
    bool timing = false;
    std::vector<hipEvent_t> events;   // pairs: start, end
    std::vector<int> event_stage;     // stage of each pair
    size_t events_used = 0;           // pairs recorded
    int densify_calls = 0;

    // debug
    float* dbg = nullptr;
    size_t dbg_cap = 0;
};

namespace pmp {

int set_error(pmp_ctx* ctx, int code, const std::string& msg);
// Stage timing: returns the pair index to close with timing_end (or -1 when off).
enum { kStageInit = 0, kStagePropagate = 1, kStageFinalize = 2 };
int timing_begin(pmp_ctx* ctx, int stage);
void timing_end(pmp_ctx* ctx, int pair);
int check_hip(pmp_ctx* ctx, hipError_t e, const char* what);
int ensure_workspace(pmp_ctx* ctx, int W, int H);
int ensure_tables(pmp_ctx* ctx, int W, int H, float sigma_color);
void host_tables(int W, int H, std::vector<float2>& rowtab, std::vector<float2>& coltab);
void derive_poses(const double* ref_pose, int N, const double* view_poses, float (*M)[9], float (*c)[3]);

}  // namespace pmp

#define PMP_HIP(ctx, expr)                                          \
    do {                                                            \
        hipError_t e_ = (expr);                                     \
        if (e_ != hipSuccess) return pmp::check_hip(ctx, e_, #expr); \
    } while (0)
