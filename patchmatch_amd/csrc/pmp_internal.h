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
constexpr int kMaxHalf = 10;   // largest patch half-width (patch_half)

// Kernel argument block of one keyframe: everything uniform across the grid.
// Passed by value (kernarg segment, read with scalar loads).
struct KfArgs {
    int W, H, N, half, stride, T;
    int tc, kc, klast;                    // centre tap index and offset ((kc, kc), the shift tap), last tap offset
    int agg_mode, topk, edge_mode, coverage, init_mode, n_refine;
    const uint8_t* ref;                   // H*W
    const uint32_t* vq[kMaxViews];        // per view: H*(W+2) packed column-pair halfwords (k_pack)
    const float2* rowtab;                 // H: (cos lat, sin lat)
    const float2* coltab;                 // W: (sin lon, cos lon)
    const float* Wc;                      // 256: exp(-a^2 / (2 sigma_color^2))
    float Ws[kMaxTaps];                   // exp(-|k|^2 / (2 sigma_space^2))
    float M[kMaxViews][9];                // R_n R_ref^T
    float c[kMaxViews][3];                // t_n - M t_ref
    float ku, hu, kv, hv;                 // u = lon*ku + hu, v = asin*kv + hv
    float umax, vmax;                     // sample clamps: W (wrap) or W-1, and H-1
    float var_floor, fold_agree, inv_k;
    float rn[kMaxViews];                  // 1/(v+1), rounded once (pooled moments of PMP_AGG_REF_FOLD)
    float depth_penalty, plane_penalty;
    float dr[kMaxRefine], nr[kMaxRefine];
    float min_depth, max_depth, inv_lo, inv_hi;
    uint32_t key0, key1;
};

// Device maps of one keyframe (grown on demand, kept).
struct SlotBuf {
    size_t cap = 0;
    int W = 0, H = 0;
    float4 *plane = nullptr, *iplane = nullptr, *oplane = nullptr;   // working, after init, final
    float *depth = nullptr, *score = nullptr, *idepth = nullptr, *iscore = nullptr;
    float *odepth = nullptr, *oscore = nullptr, *tmp0 = nullptr, *tmp1 = nullptr;
    int32_t* onghbr = nullptr;
    static constexpr size_t kBytesPerPixel = 3 * 16 + 9 * 4;
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
    uint32_t* quads = nullptr;        // H*(W+2) column-pair halfwords p(x,y) | p(x,y+1) << 8 (k_pack), built on first use
    int quads_wrap = -1;              // columns past the last: 1 wrap to 0, 1; 0 clamp; -1 not built
    uint64_t last_use = 0;            // pool clock of the last request that read it (LRU)
    int pins = 0;                     // requests in flight that read it (never evicted while > 0)
    // device bytes: the grey image and, once built, its packed halfwords
    size_t pool_bytes() const { return bytes + (quads ? quad_bytes(W, H) : 0); }
    // packed halfwords of a W x H image: rows of W + 2, the extra columns
    // x = W, W+1 repeating columns 0, 1 (wrap) or column W-1 (clamp); rounded
    // up to whole 32-bit words
    static size_t quad_bytes(int W, int H) { return ((size_t)(W + 2) * H * 2 + 3) / 4 * 4; }
};

struct QueueState;   // pmp_queue.h
struct PreState;     // pmp_preinit.hip

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

    // workspace of the one-keyframe path (grown on demand, never freed per call)
    pmp::SlotBuf ws;
    uint64_t pool_clock = 0;          // advances once per request that reads the pool
    uint8_t* mask = nullptr;          // staging of a host mask (ws.cap bytes)
    uint8_t* mask_pin = nullptr;      // pinned copy of the last host mask uploaded into mask
    size_t mask_pin_cap = 0;
    size_t mask_held = 0;             // bytes of that mask that mask holds (0: none)
    hipEvent_t mask_copied = nullptr; // recorded after the upload out of mask_pin
    bool mask_pending = false;
    std::unordered_map<int, float2*> rowtabs;   // by H: (cos lat, sin lat)
    std::unordered_map<int, float2*> coltabs;   // by W: (sin lon, cos lon)
    std::unordered_map<uint32_t, float*> wcs;   // by the bits of sigma_color: exp(-a^2 / (2 sigma^2)), a < 256
    float last_sigma = 50.0f;                   // sigma_color of the last estimator call (used by the Cleaner)
    unsigned long long* counters = nullptr;   // [0] evals, [1] candidates
    bool tables_ok = false;

    // cleaner
    float* clean_maps = nullptr;
    size_t clean_cap = 0;
    float* normals = nullptr;           // pmp_get_normals output staging, 3 floats per pixel
    size_t normals_cap = 0;
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

    // descriptors, launch configuration and graphs of the work-queue kernels
    pmp::QueueState* queue = nullptr;
    // pre-initialisation candidates and the memorised previous result
    pmp::PreState* pre = nullptr;

    // thumbnail staging (pmp_thumb.hip)
    struct {
        uint8_t* src = nullptr;
        size_t src_cap = 0;
        uint8_t* gray = nullptr;
        size_t gray_cap = 0;
        uint8_t* dst = nullptr;
        size_t dst_cap = 0;
        uint8_t* taps = nullptr;
        size_t taps_cap = 0;
    } thumb;

    // debug
    float* dbg = nullptr;
    size_t dbg_cap = 0;
    // pass epochs of the propagation (pmp_debug_epochs): mode, device array
    // (per slot of the last armed call: its tiles' epochs, then its violation
    // count), its capacity, and that call's tiles per slot and total words
    int epoch_mode = 0;
    unsigned int* epochs = nullptr;
    size_t epochs_cap = 0;
    std::vector<int> epoch_tiles;
    size_t epoch_words = 0;

    // pinned staging of host uploads (upload_host_rows)
    uint8_t* stage = nullptr;
    size_t stage_cap = 0;
};

namespace pmp {

int set_error(pmp_ctx* ctx, int code, const std::string& msg);
// Stage timing: returns the pair index to close with timing_end (or -1 when off).
enum { kStageInit = 0, kStagePropagate = 1, kStageFinalize = 2 };
int timing_begin(pmp_ctx* ctx, int stage);
void timing_end(pmp_ctx* ctx, int pair);
int check_hip(pmp_ctx* ctx, hipError_t e, const char* what);
// `rows` host rows of `row` bytes, `pitch` bytes apart, to contiguous device
// memory at dst on stream s, and waits for the copy.  The rows are packed into
// a pinned staging buffer of the context (grown on demand) and sent in one
// copy: a 2D copy from pageable memory moves row by row, about 9 us a row.
int upload_host_rows(pmp_ctx* ctx, void* dst, const uint8_t* src, size_t row, size_t rows, size_t pitch,
                     hipStream_t s);
int ensure_workspace(pmp_ctx* ctx, int W, int H);
struct Tables {
    const float2* row = nullptr;
    const float2* col = nullptr;
    const float* wc = nullptr;
};
// Bearing tables of a W x H image and the colour-weight table of sigma_color
// (built once per size / sigma, kept for the context's lifetime).
int get_tables(pmp_ctx* ctx, int W, int H, float sigma_color, Tables& t);
// Packed column pairs of a pooled image for the given column wrap (built on
// first use, on `stream`).
int ensure_quads(pmp_ctx* ctx, Image& im, int wrap, hipStream_t stream);
void host_tables(int W, int H, std::vector<float2>& rowtab, std::vector<float2>& coltab);
// One entry of the bearing tables: row i of H (cos lat, sin lat), column j of
// W (sin lon, cos lon), in double rounded once (ComputeBearingsEquirect_,
// PatchMatchCuda.cu:191-198).
float2 table_row(int i, int H);
float2 table_col(int j, int W);
void derive_poses(const double* ref_pose, int N, const double* view_poses, float (*M)[9], float (*c)[3]);
int check_params(pmp_ctx* ctx, const pmp_params* p);
void queue_free(pmp_ctx* ctx);
void preinit_free(pmp_ctx* ctx);
int ensure_slot(pmp_ctx* ctx, SlotBuf& b, int W, int H);
void free_slot(SlotBuf& b);
int build_args(pmp_ctx* ctx, const pmp_params* p, uint64_t ref_id, int N, const uint64_t* view_ids,
               const double* ref_pose, const double* view_poses, hipStream_t stream, KfArgs& a);

}  // namespace pmp

#define PMP_HIP(ctx, expr)                                          \
    do {                                                            \
        hipError_t e_ = (expr);                                     \
        if (e_ != hipSuccess) return pmp::check_hip(ctx, e_, #expr); \
    } while (0)

namespace pmp {

// The device of a context made current on the calling thread for the length
// of a call, and the caller's own current device put back when it returns:
// allocations, copies and launches land on the context's device from any
// thread, and contexts on different GPUs can be driven from one thread.
class DeviceScope {
  public:
    explicit DeviceScope(int device) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess) return;
        if (cur == device) { ok_ = true; return; }
        if (hipSetDevice(device) != hipSuccess) return;
        prev_ = cur;
        ok_ = true;
    }
    ~DeviceScope() {
        if (prev_ >= 0) (void)hipSetDevice(prev_);
    }
    DeviceScope(const DeviceScope&) = delete;
    DeviceScope& operator=(const DeviceScope&) = delete;
    bool ok() const { return ok_; }

  private:
    int prev_ = -1;
    bool ok_ = false;
};

}  // namespace pmp

// First statement of an entry point after its argument checks.
#define PMP_ON_DEVICE(ctx)                                                  \
    pmp::DeviceScope pmp_device_scope_((ctx)->device);                      \
    if (!pmp_device_scope_.ok())                                            \
        return pmp::set_error((ctx), PMP_EDEVICE, "cannot make the context's device current")
