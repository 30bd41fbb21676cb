// pmp_queue.h -- descriptors and host-side launch helpers of the work-queue
// kernels (pmp_queue.hip).  Every densification path launches these kernels:
// the one-keyframe API (a set with one slot), the batched form (up to
// PMP_MAX_BATCH slots per launch) and the server (the slots in flight, each at
// its own stage).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <tuple>
#include <utility>
#include <vector>

#include "pmp_internal.h"

namespace pmp {

constexpr int kMaxSlots = PMP_MAX_BATCH;
constexpr int kMaxPasses = 2 * 64;

// Everything a kernel needs about one keyframe slot.
struct SlotDesc {
    KfArgs a;
    float4* plane;            // working maps
    float* depth;
    float* score;
    float4* iplane;           // after RandomInitialization (copy-back outside the window)
    float* idepth;
    float* iscore;
    float4* oplane;           // final maps
    float* odepth;
    float* oscore;
    int32_t* onghbr;
    float* tmp0;              // median scratch
    float* tmp1;
    const uint8_t* mask;      // nullable
    double rel_thr;           // PostProcess median test (a double, depthmap.cc:1190)
    float min_var, inv_T, min_score;
    int postprocess;
    int pass0;                // red-black pass of this slot at launch pass 0
    int nbx_init;             // 64-pixel tiles per row
    int nbx_prop;             // propagation tiles per tile row (kTileSame same-colour pixels each)
    int nby_prop;             // propagation tile rows
    // nullable (debug): per propagation tile, the number of red-black passes
    // it has completed; epochs[nbx_prop * nby_prop] counts violations
    unsigned int* epochs;
};

// Items of one stage: the slots taking part and the prefix sums of their
// item counts (tiles or rows).
struct WorkList {
    int n;
    int total;
    int slot[kMaxSlots];
    int end[kMaxSlots];
};

struct BatchHdr {
    WorkList init, prop, rows;
};

// A descriptor set in device memory (read by the kernels) with its pinned
// host staging copy, the per-pass band heads of the propagation queue, and
// the hipGraphs captured over this set's device pointers.
using GraphKey = std::tuple<int, int, int, int, int, int>;   // passes, zero, rows, blocks, bands, postprocess
struct DescSet {
    int cap = 0;
    SlotDesc* h_slots = nullptr;
    BatchHdr* h_hdr = nullptr;
    SlotDesc* d_slots = nullptr;
    BatchHdr* d_hdr = nullptr;
    unsigned int* d_heads = nullptr;     // 8 per pass
    hipEvent_t uploaded = nullptr;       // recorded after the last upload
    bool pending = false;
    std::map<GraphKey, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
};

// Launch shape of the propagation kernel.
struct QueueRun {
    bool zero = false;        // PMP_EDGE_ZERO sampling
    bool post = false;        // any slot post-processes
    int rows = 8;             // tile rows
    int blocks = 0;           // persistent grid
    int bands = 1;            // per-XCD bands
};

// Default launch shape: 8 x 64 tiles (256-thread blocks), 3 blocks per CU.
// Measured against 16 x 64 tiles at 2 blocks per CU: 5 % faster on a
// 1920x960 keyframe with 6 views, 3 % on 960x480 with 3 views, 1 % on
// 2880x1440 with 10 views (4 blocks per CU were 1-2 % faster still on the
// last and with zero edges, 2-4 % slower on the first two).
constexpr pmp_batch_config kDefaultConfig = {4, 3, 1, 1};

struct QueueState {
    pmp_batch_config cfg = kDefaultConfig;
    int cus = 0;
    DescSet plain;            // the one-keyframe API
    DescSet batch;            // pmp_batch_*
    std::vector<SlotBuf> bufs;
    int n_last = 0;
};

int queue_state(pmp_ctx* ctx, QueueState*& out);
// Landmark seeding of one slot of an uploaded descriptor set (pmp_preinit.hip).
int preinit_landmarks_slot(pmp_ctx* ctx, const SlotDesc* slot_dev, int W, int H, const double* pose, int64_t n,
                           const double* xyz, bool zero);
// Warm start of one keyframe slot of a descriptor set (the served form): the
// previous keyframe's final plane4 / depth maps (W x H pixels, pinned host
// memory valid until the stream has passed this stage) and its pose, warped
// into the slot's maps as pmp_preinit_warp does.  Stream-ordered on
// ctx->stream, no host synchronisation (scratch growth aside).
int preinit_warp_slot(pmp_ctx* ctx, const SlotDesc* slot_dev, int W, int H, const double* pose,
                      const double* prev_pose, const float* prev_plane4, const float* prev_depth, bool zero);
int desc_alloc(pmp_ctx* ctx, DescSet& d, int cap);
void desc_free(DescSet& d);
// Waits until the staging copy of d may be rewritten (its previous upload has executed).
int desc_begin(pmp_ctx* ctx, DescSet& d);
int desc_upload(pmp_ctx* ctx, DescSet& d, int n, hipStream_t stream);
// Descriptor of one keyframe over the maps of b.  dmask (device, nullable)
// replaces kf.mask_id when given.  items receives the keyframe's init tiles,
// propagation tiles (for `rows` tile rows) and rows.
int slot_fill(pmp_ctx* ctx, const pmp_params* p, const pmp_keyframe& kf, const SlotBuf& b, const uint8_t* dmask,
              int pass0, int rows, hipStream_t stream, SlotDesc& d, int items[3]);
void fill_list(WorkList& L, int n, const int* items);
int queue_run(pmp_ctx* ctx, const QueueState& q, const pmp_batch_config& c, bool zero, bool post, QueueRun& r);
int q_clear(pmp_ctx* ctx, hipStream_t s, DescSet& d, int passes);
int q_init(pmp_ctx* ctx, hipStream_t s, DescSet& d, const QueueRun& r);
int q_score(pmp_ctx* ctx, hipStream_t s, DescSet& d, const QueueRun& r);
int q_prop(pmp_ctx* ctx, hipStream_t s, DescSet& d, const QueueRun& r, int pass);
int q_finalize(pmp_ctx* ctx, hipStream_t s, DescSet& d, const QueueRun& r);
// clear + init + `passes` passes + finalize, replayed from a graph cached in d
// when use_graph (the set's contents may change between replays).
int q_densify(pmp_ctx* ctx, hipStream_t s, DescSet& d, const QueueRun& r, int passes, bool use_graph);
// n words of p set to v with a fill kernel (capture-safe).
int q_fill_u32(pmp_ctx* ctx, hipStream_t s, unsigned int* p, size_t n, unsigned int v);
// With pmp_debug_epochs on: gives each of the n slots a segment of the
// context's epoch array (tiles[k] epochs set to pass0, then a zero violation
// count; mode 2 starts slot 0's tile 0 one pass ahead), on stream s, and
// points slots[k].epochs at it.  Off: leaves the slots' epochs null.
int epochs_arm(pmp_ctx* ctx, hipStream_t s, int n, SlotDesc* slots, const int* tiles, int pass0);

// The one-keyframe API's current keyframe as a one-slot set over the context
// workspace, uploaded (pmp_estimator.hip).  dmask (device, nullable) is the
// mask; with_mask: use the pmp_set_mask image when dmask is null.
int plain_prepare(pmp_ctx* ctx, const pmp_params* p, const uint8_t* dmask, bool with_mask, int pass0, bool propagates,
                  QueueState*& q, QueueRun& r);
// A host mask (nullable) copied to the workspace's staging buffer.
int stage_mask(pmp_ctx* ctx, const uint8_t* mask_host, const uint8_t*& dmask);

}  // namespace pmp
