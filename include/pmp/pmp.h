/*
 * pmp.h -- C-ABI of the MI355X-native PatchMatch-Stereo-Panorama densifier.
 *
 * Plain C linkage, POD arguments, no exceptions across the boundary.  Every
 * entry point returns PMP_OK (0) or a negative PMP_E* status; the reason is
 * available from pmp_last_error().  One context per host thread (contexts
 * are not internally locked); a context owns its device memory, its HIP
 * stream and its image pool.  The caller owns all host buffers.
 *
 * Each entry point names the reference interface it replaces (paths are in
 * the RoblabWh/PatchMatch tree).  The reference's CUDA boundary is the set
 * of extern "C" declarations at src/openvslam/depthmap.cc:11-21, defined in
 * src/openvslam/PatchMatchCuda.cu:419-912, driven by
 * DepthmapEstimator::ComputePatchMatch (depthmap.cc:541-806); its C++
 * boundary is dense::DepthmapEstimator / DepthmapCleaner / DepthmapPruner
 * (depthmap.h:122-291).
 *
 * Conventions:
 *   images   uint8 grey, equirectangular, row pitch in bytes;
 *   poses    world->camera rotation (row-major) and translation, doubles
 *            (data/keyframe.cc:235-243: rot_cw, trans_cw);
 *   planes   p with p.X = 1 for camera-frame points X on the plane, so the
 *            depth along the unit bearing b is 1/|b.p| (PatchMatchCuda.cu:31-43);
 *   maps     H*W, row-major; plane maps as 4 floats per pixel (x, y, z, 0).
 */
#ifndef PMP_PMP_H
#define PMP_PMP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PMP_MAX_VIEWS 10

enum {
    PMP_OK = 0,
    PMP_EINVAL = -1,     /* bad argument */
    PMP_ENOMEM = -2,     /* device or host allocation failed */
    PMP_EDEVICE = -3,    /* HIP runtime error */
    PMP_ENOTFOUND = -4,  /* unknown image id */
    PMP_ESTATE = -5,     /* call out of order (e.g. no views set) */
    PMP_EBUSY = -6,      /* not finished within the time limit */
    PMP_ECANCELED = -7   /* a served request ended by pmp_server_cancel or pmp_server_reset */
};

/* PMP_AGG_REF_FOLD: the fold of PatchMatchCuda.cu:161-171 with the moment sums
 * pooled across views, as there (zeroed once at :57-62, outside the view loop
 * :101), so view v's NCC is taken over views 0..v together;
 * PMP_AGG_MAX: maximum of per-view NCCs (ComputePlaneScore, depthmap.cc:1104-1127);
 * PMP_AGG_TOPK_MEAN: mean of the topk largest per-view NCCs;
 * PMP_AGG_FOLD_INDEP: the fold over independent per-view NCCs (no counterpart
 * in the reference). */
enum { PMP_AGG_REF_FOLD = 0, PMP_AGG_MAX = 1, PMP_AGG_TOPK_MEAN = 2, PMP_AGG_FOLD_INDEP = 3 };
enum { PMP_EDGE_WRAP = 0, PMP_EDGE_CLAMP = 1, PMP_EDGE_ZERO = 2 };
enum { PMP_COVER_REF_GRID = 0, PMP_COVER_FULL = 1 };
enum { PMP_INIT_REF_UNIFORM = 0, PMP_INIT_INVDEPTH_HEMI = 1 };
enum { PMP_MAPS_WORK = 0, PMP_MAPS_INIT = 1, PMP_MAPS_FINAL = 2 };
enum { PMP_COLOR_GRAY = 0, PMP_COLOR_RGB = 1, PMP_COLOR_BGR = 2 };   /* camera::color_order_t */

typedef struct pmp_ctx pmp_ctx;

/* Every constant the reference hard-codes, with the reference value as the
 * default (pmp_default_params).  Layout is part of the ABI. */
typedef struct pmp_params {
    int32_t patch_half;        /* 5: patch_size_-2 (depthmap.cc:666,725), taps -half..half */
    int32_t patch_stride;      /* 2: PatchMatchCuda.cu:104-106 */
    int32_t iterations;        /* red-black iterations (Dense.Iterations) */
    int32_t n_refine;          /* 6 random refinements per pass (:363-395), <= 8 */
    int32_t agg_mode;          /* PMP_AGG_*: view aggregation (:161-171 pooled fold) */
    int32_t topk;              /* k of PMP_AGG_TOPK_MEAN */
    int32_t edge_mode;         /* PMP_EDGE_*: neighbour sampling outside the image */
    int32_t coverage_mode;     /* PMP_COVER_*: reference launch grids or every interior pixel */
    int32_t init_mode;         /* PMP_INIT_*: random hypothesis distribution */
    int32_t postprocess;       /* 1: run PostProcess (depthmap.cc:1169-1208) */
    uint64_t seed;             /* key of the Philox streams */
    float min_depth, max_depth;            /* init range and PostProcess range */
    float init_inv_depth_min, init_inv_depth_max; /* PMP_INIT_INVDEPTH_HEMI range (1/m) */
    float min_score;           /* 0.1: PostProcess score threshold */
    float sigma_color;         /* 50 */
    float sigma_space;         /* 5 */
    float var_floor;           /* 0.1: NCC variance floor */
    float fold_agree;          /* 0.25: fold agreement */
    float depth_penalty;       /* 1250 */
    float plane_penalty;       /* 8 */
    float refine_depth_range;  /* 0.02 */
    float refine_normal_range; /* 0.5 */
    float refine_depth_decay;  /* 0.3 */
    float refine_normal_decay; /* 0.8 */
    float min_patch_sd;        /* 0: low-texture gate (depthmap.cc:739) */
    double median_rel_thr;     /* 0.05: PostProcess median test, a double as the literal
                                  of depthmap.cc:1190 (the float ratio is compared with it) */
} pmp_params;

typedef struct pmp_pose {
    double R[9];   /* world->camera rotation, row-major */
    double t[3];   /* world->camera translation */
} pmp_pose;

/* ---- context ----------------------------------------------------------- */

/* Replaces the per-call cudaMalloc/cudaFree of DepthmapEstimator
 * (depthmap.cc:522,557-558,652-656,771-800): buffers live in the context and
 * grow on demand.  device: HIP device ordinal.  Every call on a context (and
 * on a server) runs on that device from any thread and leaves the calling
 * thread's current HIP device as it found it. */
int pmp_create(pmp_ctx** out, int32_t device);
void pmp_destroy(pmp_ctx* ctx);
/* Replaces printing cudaGetErrorString and continuing (PatchMatchCuda.cu:447-451). */
const char* pmp_last_error(const pmp_ctx* ctx);
/* Reference operating values. */
void pmp_default_params(pmp_params* p);
/* Waits for the context's stream. */
int pmp_synchronize(pmp_ctx* ctx);

/* ---- images ------------------------------------------------------------ */

/* Upload (or replace) a grey image in the context's pool under id.
 * Replaces InitTextureImages (PatchMatchCuda.cu:697-797, float textures built
 * from depthmap.cc:569-571); images stay resident as uint8 plus a packed
 * 2x2-neighbourhood copy for bilinear sampling.  on_device != 0: pixels is a
 * device pointer on the context's device.  Host rows (pitch bytes apart) go
 * through a pinned staging buffer of the context in one copy; the call
 * returns once the image is on the device. */
int pmp_image_put(pmp_ctx* ctx, uint64_t id, const uint8_t* pixels, int32_t width, int32_t height,
                  int64_t pitch, int32_t on_device);
int pmp_image_drop(pmp_ctx* ctx, uint64_t id);
/* Copy a pooled image (width*height bytes, rows packed) to host memory. */
int pmp_image_get(pmp_ctx* ctx, uint64_t id, uint8_t* pixels);

/* The frame thumbnails of tracking_module.cc:82-93 and data/frame.cc:53-54,
 * :136-138, on the device: a full-resolution camera image (channels 1, 3 or 4
 * in color_order) converted to grey (cv::cvtColor) at full resolution, then
 * resized with INTER_LINEAR (cv::resize) to out_width x out_height and pooled
 * under id (replacing it).  8-bit fixed point as OpenCV's published formulas.
 * on_device != 0: pixels is a device pointer. */
int pmp_image_put_thumbnail(pmp_ctx* ctx, uint64_t id, const uint8_t* pixels, int32_t width, int32_t height,
                            int64_t pitch, int32_t channels, int32_t color_order, int32_t out_width,
                            int32_t out_height, int32_t on_device);
/* cv::resize INTER_LINEAR of an 8-bit image with 1, 3 or 4 interleaved
 * channels (the colour and mask thumbnails, frame.cc:53, :136-138); src and
 * dst both host or both device (on_device); dst rows are out_width*channels. */
int pmp_resize(pmp_ctx* ctx, const uint8_t* src, int32_t width, int32_t height, int64_t pitch, int32_t channels,
               uint8_t* dst, int32_t out_width, int32_t out_height, int32_t on_device);
/* Number of images and device bytes held by the pool. */
int pmp_image_pool_stats(const pmp_ctx* ctx, int64_t* count, int64_t* bytes);
/* Width and height of a pooled image (PMP_ENOTFOUND if absent): lets a caller
 * size the buffers of pmp_image_get and of the masks it passes.  No
 * counterpart in the reference, whose images are cv::Mat with their size. */
int pmp_image_size(const pmp_ctx* ctx, uint64_t id, int32_t* width, int32_t* height);

/* ---- one keyframe (DepthmapEstimator, depthmap.cc:541-806) -------------- */

/* Choose the reference image and its neighbours for the next calls.
 * Replaces InitConstantMem (PatchMatchCuda.cu:885-912) and the pose algebra
 * of depthmap.cc:577-649 (ref pose = rot_cw/trans_cw of the centre keyframe,
 * neighbour poses = their own rot_cw/trans_cw). */
int pmp_set_views(pmp_ctx* ctx, uint64_t ref_id, const pmp_pose* ref_pose, int32_t n_views,
                  const uint64_t* view_ids, const pmp_pose* view_poses);

/* RandomInitialization (PatchMatchCuda.cu:656-663): the memsets of
 * depthmap.cc:657-659, random depth/plane per pixel and the initial score.
 * Also snapshots the maps for pmp_finalize (the reference keeps them in the
 * host result, depthmap.cc:673-687).  Replaces ComputeRandomSeed too: the
 * random streams are counter-based, keyed on params->seed. */
int pmp_random_initialization(pmp_ctx* ctx, const pmp_params* params);

/* ScorePlaneDepth (PatchMatchCuda.cu:528-565): score the working plane map
 * at every pixel of the scoring grid (params->coverage_mode). */
int pmp_score_plane_depth(pmp_ctx* ctx, const pmp_params* params);

/* PatchMatchRedBlackPass (PatchMatchCuda.cu:568-615): iterations
 * [it_begin, it_end), each a pass over (i+j) even pixels then (i+j) odd. */
int pmp_red_black_pass(pmp_ctx* ctx, const pmp_params* params, int32_t it_begin, int32_t it_end);

/* Select a pooled image (same size as the reference) as the device-resident
 * mask used by pmp_finalize/pmp_densify when their mask argument is NULL;
 * mask_id < 0 clears it (then NULL means every pixel is valid). */
int pmp_set_mask(pmp_ctx* ctx, int64_t mask_id);

/* Copy-back window, low-texture gate and PostProcess of ComputePatchMatch
 * (depthmap.cc:673-687, :731-767, :1169-1208).  mask: host H*W bytes,
 * 0 = ignored pixel; NULL = the pmp_set_mask image, or all valid. */
int pmp_finalize(pmp_ctx* ctx, const pmp_params* params, const uint8_t* mask);

/* ---- pre-initialisation (depthmap.cc:690-691), between
 * pmp_random_initialization and pmp_red_black_pass.  Candidates are scored
 * with params' scorer at scored pixels; a candidate replaces a pixel's
 * plane/depth/score (working maps and the copy-back snapshot) when it has the
 * highest score of the pixel's candidates (the first in loop order among
 * equal ones) and beats the pixel's score: the sequential outcome of the
 * reference's OpenMP loops.  mask: host H*W bytes (0 = ignored), NULL = the
 * pmp_set_mask image or none.  accepted (nullable): pixels changed; asking
 * for it waits for the device. */

/* PreInitSparsePCL2 (depthmap.cc:361-432): n landmarks (world xyz, doubles)
 * seed the pixel each reprojects to with the camera-facing plane at its
 * distance. */
int pmp_preinit_landmarks(pmp_ctx* ctx, const pmp_params* params, int64_t n, const double* xyz,
                          const uint8_t* mask, int64_t* accepted);
/* MemorizeResult (depthmap.cc:808-829): keep the final plane and depth maps
 * and the reference pose of the last finalize for pmp_preinit_warp. */
int pmp_memorize_result(pmp_ctx* ctx);
/* Drop the memorised result (DepthmapEstimator::reset, depthmap.cc:261-270):
 * pmp_preinit_warp is a no-op until the next pmp_memorize_result. */
int pmp_forget_result(pmp_ctx* ctx);
/* PreInitResultWarp (depthmap.cc:434-516): move every pixel of the memorised
 * result with depth >= 1e-4 along its plane into the current reference; the
 * candidate's normal is the memorised plane's direction, as in the reference
 * (not rotated into the current frame).  No-op before the first
 * pmp_memorize_result. */
int pmp_preinit_warp(pmp_ctx* ctx, const pmp_params* params, const uint8_t* mask, int64_t* accepted);

/* The whole estimator: init, params->iterations red-black iterations,
 * finalize.  Replaces the device part of DepthmapEstimator::ComputePatchMatch
 * (depthmap.cc:555-767). */
int pmp_densify(pmp_ctx* ctx, const pmp_params* params, const uint8_t* mask);

/* Upload working maps (replaces the H->D copies of depthmap.cc:717-719).
 * plane4: H*W*4 floats; any pointer may be NULL (left unchanged). */
int pmp_set_maps(pmp_ctx* ctx, const float* plane4, const float* depth, const float* score);

/* Download maps (replaces the D->H copies of depthmap.cc:669-671, :727-729).
 * which: PMP_MAPS_WORK (working maps), PMP_MAPS_INIT (after initialisation)
 * or PMP_MAPS_FINAL (DepthmapEstimatorResult: depth, plane, score, nghbr,
 * depthmap.h:71-79).  plane gets 4 floats per pixel; NULL pointers skipped;
 * nghbr only for PMP_MAPS_FINAL. */
int pmp_get_maps(pmp_ctx* ctx, int32_t which, float* plane4, float* depth, float* score, int32_t* nghbr);

/* Normals of a map set (the plane of DepthmapEstimatorResult, depthmap.h:71-79,
 * as a normal map): H*W*3 floats, the unit normal -p/|p| of the plane p = n/d.
 * Where the plane lies in front of the camera (p . b = 1/depth > 0) it faces
 * the camera.  (0, 0, 0) where p = 0.  Computed on the device. */
int pmp_get_normals(pmp_ctx* ctx, int32_t which, float* normal3);

/* Counters of the last pmp_red_black_pass/pmp_densify: hypotheses scored
 * (evals) and scored+pre-rejected (candidates), summed over pixels. */
int pmp_eval_counts(pmp_ctx* ctx, int64_t* evals, int64_t* candidates);

/* Device time per stage, from HIP events recorded on the context's stream
 * around every launch while timing is enabled (off by default).  Replaces the
 * cudaEvent + printf timing of every reference wrapper
 * (PatchMatchCuda.cu:421-441, :570-601).  pmp_get_timing waits for the
 * stream, reports the launches recorded since the previous call, and resets. */
typedef struct pmp_timing {
    double init_ms;           /* random initialisation + initial score */
    double propagate_ms;      /* sum over red-black pass launches */
    double finalize_ms;       /* copy-back, gate, PostProcess */
    int32_t propagate_launches;
    int32_t densify_calls;
} pmp_timing;
int pmp_set_timing(pmp_ctx* ctx, int32_t enable);
int pmp_get_timing(pmp_ctx* ctx, pmp_timing* out);

/* ---- many keyframes per launch (batched form) --------------------------- */

#define PMP_MAX_BATCH 128

/* One keyframe of a batch: pooled reference and neighbour images with their
 * world->camera poses (the arguments of pmp_set_views), an optional pooled
 * mask, and the keyframe's depth range.  The reference derives the range per
 * keyframe from its landmarks (depthmap_computation_module.cc:123-131);
 * min_depth/max_depth > 0 replace the params' values for this keyframe. */
typedef struct pmp_keyframe {
    uint64_t ref_id;
    pmp_pose ref_pose;
    int32_t n_views;
    int32_t reserved;
    uint64_t view_ids[PMP_MAX_VIEWS];
    pmp_pose view_poses[PMP_MAX_VIEWS];
    int64_t mask_id;          /* pooled mask image (0 = ignored pixel), -1: none */
    float min_depth, max_depth;
} pmp_keyframe;

/* Result-invariant launch configuration of the batched kernels: every value
 * gives bit-identical maps, only the speed differs. */
typedef struct pmp_batch_config {
    int32_t tile_rows;        /* rows per propagation tile (2, 4 or 8); the block has 64*tile_rows threads */
    int32_t blocks_per_cu;    /* persistent propagation blocks per CU (0: as many as fit) */
    int32_t use_graph;        /* 1: replay the whole batch (init, 2*iterations passes, finalize) as one hipGraph */
    int32_t xcd_bands;        /* 1: each XCD drains its own contiguous band of tiles first */
} pmp_batch_config;

/* Densify n <= PMP_MAX_BATCH keyframes of any sizes and view counts with one
 * set of params: one work-queue launch per stage walks the (keyframe, tile)
 * items of every keyframe.  Per keyframe the maps are bit-identical to
 * pmp_set_views + pmp_densify with the same params.  Asynchronous on the
 * context stream; results stay on the device until pmp_batch_get_maps. */
int pmp_batch_densify(pmp_ctx* ctx, const pmp_params* params, int32_t n, const pmp_keyframe* kfs);
/* Final maps of keyframe k of the last batch (host pointers, nullable):
 * plane 4 floats per pixel, depth, score, nghbr (as PMP_MAPS_FINAL). */
int pmp_batch_get_maps(pmp_ctx* ctx, int32_t k, float* plane4, float* depth, float* score, int32_t* nghbr);
int pmp_batch_set_config(pmp_ctx* ctx, const pmp_batch_config* cfg);
int pmp_batch_get_config(pmp_ctx* ctx, pmp_batch_config* cfg);
/* Times the red-black passes of the batch (init, then every pass, from the
 * same starting maps) under every configuration in the search space (tile
 * rows x blocks per CU x XCD bands, at least 8 waves per CU), keeps the
 * fastest as the context's configuration and writes it to *best (nullable).
 * The batch's results are not kept (run pmp_batch_densify afterwards). */
int pmp_batch_autotune(pmp_ctx* ctx, const pmp_params* params, int32_t n, const pmp_keyframe* kfs,
                       pmp_batch_config* best);

/* ---- served form: a request stream with continuous batching ------------- */

/* The reference densifies one keyframe at a time on the densify thread
 * (depthmap_computation_module.cc:374-413, one synchronous ComputePatchMatch
 * per keyframe).  A server owns a device context, an on-device LRU image pool
 * admitted against an HBM budget, and a scheduler thread: every scheduler
 * step launches one stage (init, one red-black pass, or finalize) of every
 * keyframe in flight, each at its own stage, so requests that arrive while
 * others run join at the next step.  Per request the maps are bit-identical
 * to pmp_densify with the same params. */
typedef struct pmp_server pmp_server;

typedef struct pmp_server_opts {
    int32_t device;
    int32_t max_slots;          /* keyframes in flight, 1..PMP_MAX_BATCH (0: 16) */
    int64_t hbm_budget;         /* device bytes for pooled images (5 B/pixel) and keyframe maps
                                   (84 B/pixel); 0: 64 GiB */
    int32_t edge_mode;          /* PMP_EDGE_* of every request (neighbour packing is shared) */
    int32_t inflight_kpx;       /* admission budget in thousands of pixels: a queued request is
                                   admitted (oldest first) only while the keyframes in flight plus
                                   it stay within the budget, or when nothing is in flight; 0: no
                                   budget (only max_slots limits) */
    pmp_batch_config config;    /* tile_rows, blocks_per_cu, xcd_bands of the step kernels;
                                   tile_rows 0: the default (4, i.e. 8 x 64 tiles, 3 blocks per CU) */
    int32_t tail_stages;        /* latency-first admission: while at most one request is queued, it
                                   joins the requests in flight only once each of them has at most
                                   this many stages left (a request's stages: init, 2 per
                                   iteration, finalize), so that they are not slowed by sharing
                                   their pass launches; with two or more queued, admission is free
                                   (throughput first); 0: always free */
} pmp_server_opts;

/* One densification request: the keyframe (pooled ids, poses, mask id, depth
 * range) plus host pixels for every id the pool does not hold yet (read only
 * during pmp_server_submit; nullable for pooled ids), and optionally the
 * keyframe's landmarks (world positions, 3 doubles each; copied during
 * pmp_server_submit), which seed the maps after the random initialisation as
 * pmp_preinit_landmarks does (PreInitSparsePCL2, depthmap.cc:361-432). */
typedef struct pmp_request {
    pmp_keyframe kf;
    int32_t width, height;
    const uint8_t* ref_pixels;
    const uint8_t* view_pixels[PMP_MAX_VIEWS];
    const uint8_t* mask_pixels;
    int64_t n_landmarks;
    const double* landmark_xyz;
    /* Warm start (PreInitResultWarp, depthmap.cc:434-516): the previous
     * keyframe's final plane (4 floats per pixel) and depth maps, of this
     * request's size, and that keyframe's pose; after the random
     * initialisation (and the landmark seeding) each previous pixel's plane is
     * warped into this keyframe and kept where it scores better.  All three
     * null: no warm start.  Copied at submit. */
    const float* prev_plane4;
    const float* prev_depth;
    const pmp_pose* prev_pose;
} pmp_request;

typedef struct pmp_request_info {
    double submit_ms, admit_ms, done_ms;   /* server clock: ms since pmp_server_create */
    int32_t steps;                          /* scheduler steps that carried this request */
    int32_t width, height, n_views;
} pmp_request_info;

typedef struct pmp_server_info {
    int64_t submitted, completed, failed, steps;
    int64_t pool_images, pool_bytes, maps_bytes, uploads, evictions;
    int32_t in_flight, queued, max_in_flight, cancelled;
} pmp_server_info;

int pmp_server_create(pmp_server** out, const pmp_server_opts* opts);
/* Finishes every submitted request (queued ones too, unless paused), then
 * releases the device (request_terminate, :596-615). */
void pmp_server_destroy(pmp_server* srv);
/* The reason of the last failing pmp_server_* call made on this thread. */
const char* pmp_server_last_error(pmp_server* srv);
/* Queues a request with its own params (edge_mode must be the server's).
 * Missing images are uploaded (and packed) before it returns; waits while
 * the budget cannot hold them next to the images of requests in flight;
 * PMP_ENOMEM if it never could. */
int pmp_server_submit(pmp_server* srv, const pmp_request* req, const pmp_params* params, uint64_t* ticket);
/* PMP_OK once the request is done (or failed), PMP_EBUSY after timeout_ms
 * (< 0: no limit). */
int pmp_server_wait(pmp_server* srv, uint64_t ticket, int32_t timeout_ms);
/* Copies the final maps (as PMP_MAPS_FINAL; nullable pointers) and the
 * request's timings, then forgets the ticket.  A failed request returns its
 * error status. */
int pmp_server_fetch(pmp_server* srv, uint64_t ticket, float* plane4, float* depth, float* score, int32_t* nghbr,
                     pmp_request_info* info);
int pmp_server_stats(pmp_server* srv, pmp_server_info* out);
/* paused != 0: admit no queued request until resumed (requests in flight
 * continue); submissions are still accepted.  The densify thread's
 * request_pause / resume (depthmap_computation_module.cc:553-589). */
int pmp_server_pause(pmp_server* srv, int32_t paused);
/* Ends one request: a queued one at once, one in flight after the steps
 * already issued for it have run (its slot then takes the next request).
 * The ticket stays until fetched; wait returns and fetch gives
 * PMP_ECANCELED.  A request already finished keeps its result.  No
 * counterpart in the reference, whose densify thread has no per-keyframe
 * cancellation. */
int pmp_server_cancel(pmp_server* srv, uint64_t ticket);
/* request_reset (depthmap_computation_module.cc:503-546): cancels every
 * queued and in-flight request (their tickets fetch PMP_ECANCELED), waits
 * until none is in flight, and empties the image pool (the reference clears
 * its keyframe queue and depth-map database).  Blocks, as there; submissions
 * wait meanwhile.  The server stays usable. */
int pmp_server_reset(pmp_server* srv);

/* ---- geometric consistency (DepthmapCleaner, depthmap.cc:1320-1377) ----- */

/* depth_maps[0] is the map to clean (the reference swaps its centre map to
 * index 0), depth_maps[1..n_maps-1] the others; host pointers, W*H floats.
 * A pixel is kept when at least min_consistent maps (itself included) agree
 * within same_depth_thr * depth. */
int pmp_clean(pmp_ctx* ctx, int32_t width, int32_t height, int32_t n_maps, const float* const* depth_maps,
              const pmp_pose* poses, float same_depth_thr, int32_t min_consistent, float* cleaned);
/* pmp_clean with two optional tests that a view must also pass to count
 * (the reference has neither; multi-view fusion filters of this kind bound
 * the reprojection and normal disagreement):
 *   max_reproj_px > 0: the point map o holds at the reprojection (its
 *     interpolated depth along the bearing of that continuous pixel),
 *     projected back into map 0, lands within max_reproj_px pixels of the
 *     pixel being cleaned (longitude wrapped);
 *   max_normal_deg in (0, 90): the world-frame normals of the two maps (map
 *     0 at the pixel, map o at the pixel nearest the reprojection) are at
 *     most max_normal_deg apart; plane4_maps (n_maps host pointers, W*H*4
 *     floats as pmp_get_maps returns them) are then required.
 * Values <= 0 switch a test off; with both off this is pmp_clean. */
/* pmp_clean without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under graph capture,
 * tight serving loops).  Same arguments and results as pmp_clean; the staged
 * maps, the cleaned map and the bearing tables live in one file-scope device
 * array of 16 MiB per GPU (held while the library is loaded), so the call
 * makes no device allocation.  PMP_EINVAL when (n_maps + 1) * width * height
 * + 2 * (width + height) exceeds 4194304 floats (640x320 takes any n_maps up
 * to 16, 1280x640 up to 4).  Calls on one GPU share the array: calls that
 * overlap on one GPU race. */
int pmp_clean_noalloc(pmp_ctx* ctx, int32_t width, int32_t height, int32_t n_maps, const float* const* depth_maps,
                      const pmp_pose* poses, float same_depth_thr, int32_t min_consistent, float* cleaned);
int pmp_clean_ex(pmp_ctx* ctx, int32_t width, int32_t height, int32_t n_maps, const float* const* depth_maps,
                 const float* const* plane4_maps, const pmp_pose* poses, float same_depth_thr, int32_t min_consistent,
                 float max_reproj_px, float max_normal_deg, float* cleaned);

/* ---- point fusion (DepthmapPruner, depthmap.cc:1405-1696) --------------- */

/* AddView (depthmap.cc:1405-1423): append a cleaned depth map, its pose and
 * its BGR thumbnail (W*H*3 bytes) to the device-resident queue. */
int pmp_pruner_add_view(pmp_ctx* ctx, const float* cleaned_depth, const pmp_pose* pose, const uint8_t* bgr,
                        int32_t width, int32_t height);
/* Prune (depthmap.cc:1535-1696): fuse the head map against the later maps
 * (which are modified, as in the reference); results are identical to the
 * reference's raster-order loop.  Writes the pruned head map (nullable) and
 * up to cap points: pts (3 doubles, world frame), cols (3 bytes BGR), kps
 * (row, col); *count gets the number of kept points (may exceed cap). */
int pmp_pruner_prune(pmp_ctx* ctx, float same_depth_thr, float* pruned_depth, int64_t cap, double* pts,
                     uint8_t* cols, int32_t* kps, int64_t* count);
int pmp_pruner_pop_head(pmp_ctx* ctx);           /* PopHead (depthmap.cc:1429-1446) */
int pmp_pruner_reset(pmp_ctx* ctx);              /* reset (depthmap.cc:1478-1499) */
int pmp_pruner_size(const pmp_ctx* ctx);         /* queue length */
int pmp_pruner_get_map(pmp_ctx* ctx, int32_t index, float* depth);  /* read back a queued map */

/* ---- densify-thread front-end (host only; include/pmp/frontend.hpp) ------ */

/* score_view (depthmap_computation_module.cc:147-198): over the landmarks
 * common to two keyframes (matched by id; positions from cur), the fraction
 * seen from the two camera centres under an angle in (pi/60, pi/6).  NaN
 * without common landmarks (the reference's 0/0), which fails the 0.25
 * gate.  Poses: 12 doubles, world->camera R (row-major) then t. */
double pmp_score_view(const double* prev_pose, int64_t n_prev, const uint64_t* prev_ids, const double* prev_xyz,
                      const double* cur_pose, int64_t n_cur, const uint64_t* cur_ids, const double* cur_xyz);
/* compute_depth_intervall (:123-131) with computePercentileDepth_10_90
 * (data/keyframe.cc:428-451): min 1e-5, max 1.1 x the landmark distance at
 * sorted index n/10*9.  PMP_EINVAL for n = 0. */
int pmp_depth_interval(const double* pose, int64_t n, const double* xyz, double* min_depth, double* max_depth);

/* ---- output formats (host only; no context needed) ----------------------- */

/* The equirectangular camera of the map (camera/equirectangular.cc:93-103). */
typedef struct pmp_camera_info {
    const char* name;
    int32_t cols, rows;
    double fps;
} pmp_camera_info;

/* One densified keyframe: its pose and the DepthmapPrunerResult of
 * depthmap.h:243-252 (points3d world xyz, colors BGR, keypts (row, col)). */
typedef struct pmp_dense_keyframe {
    uint64_t id;
    uint64_t src_frm_id;
    double timestamp;
    pmp_pose pose;            /* world->camera */
    int64_t n_points;
    const double* points;     /* n*3 */
    const uint8_t* colors;    /* n*3, B G R */
    const int32_t* keypts;    /* n*2, row col */
} pmp_dense_keyframe;

/* save_message_pack_dense (io/map_database_io.cc:22-44): MessagePack of
 * {"cameras", "keyframes", "landmarks"} with one landmark per point
 * (dense_pointcloud_module.cc:153-174), the file export_dense_msg_to_ply.py
 * reads. */
int pmp_write_dense_msgpack(const char* path, const pmp_camera_info* camera, int32_t n_keyframes,
                            const pmp_dense_keyframe* keyframes);
/* save_point_cloud_dense (io/map_database_io.cc:46-81): ASCII PLY with
 * float x y z and uchar red green blue per vertex; rgb: n*3 (R G B). */
int pmp_write_ply(const char* path, int64_t n, const double* xyz, const uint8_t* rgb);

/* ---- diagnostics --------------------------------------------------------- */

/* Evaluate the device elementary functions (see pmp_math.h) on n inputs:
 * which 0 rcp 1 rsqrt 2 sqrt 3 asin 4 acos 5 atan2(a,b) 6 exp 7 log 8 sin2pi
 * 9 cos2pi 10 latitude of the direction (b, a, b) 11 longitude of the
 * direction (a, 0, b).  Host pointers. */
int pmp_debug_eval_math(pmp_ctx* ctx, int32_t which, int64_t n, const float* a, const float* b, float* out);
/* The standard normals of refinement for pixel (i, j), iteration it (W wide
 * image, params->seed), n <= 24. */
int pmp_debug_refine_normals(pmp_ctx* ctx, const pmp_params* params, int32_t width, int32_t i, int32_t j,
                             int32_t it, int32_t n, float* out);

/* Pass epochs of the propagation kernel (race detection for the persistent
 * work queue; no counterpart in the reference, whose passes are plain
 * launches, PatchMatchCuda.cu:273-407).  With mode 1 the red-black passes of
 * pmp_densify, pmp_red_black_pass and pmp_batch_densify (every keyframe of
 * the batch; not its autotuner) record per tile how many passes it has
 * completed, and every tile checks on entry to pass pp that it has completed
 * pp passes and that the 8 tiles around it (whose maps it reads) have
 * completed pp or pp + 1; each failed check counts one violation.  Mode 2 also
 * starts tile 0 one pass ahead, a violation forced from the host (exactly one
 * in a run that is otherwise correct).  Mode 0 turns it off.  The maps are the
 * same in every mode. */
int pmp_debug_epochs(pmp_ctx* ctx, int32_t mode);
/* The record of the last pass-epoch call: propagation tiles (over every
 * keyframe of a batch), violations, and the least and greatest number of
 * passes a tile completed (both equal to 2 x iterations after a pmp_densify
 * or pmp_batch_densify).  PMP_ESTATE before any such call.
 * Synchronises the context stream. */
int pmp_debug_epoch_report(pmp_ctx* ctx, int32_t* tiles, int32_t* violations, int32_t* min_epoch,
                           int32_t* max_epoch);

/* roctx ranges around the main calls (pmp_densify and the other stages,
 * pmp_batch_*, pmp_clean, pmp_pruner_prune, pmp_preinit_*, pmp_image_put, and
 * the served form's uploads and steps), for rocprofv3 --marker-trace: host
 * ranges of the enqueue, beside the kernel trace.  The reference brackets its
 * wrappers with cudaEvent timings printed to stdout instead
 * (PatchMatchCuda.cu:421-441).  Process-wide; also turned on by the
 * environment variable PMP_ROCTX=1.  The ROCm marker library is opened at run
 * time: PMP_ENOTFOUND when it cannot be, and tracing stays off. */
int pmp_set_roctx(int32_t on);

/* Copy `bytes` (multiple of 16) between two context-owned device buffers with
 * a streaming kernel, `repeats` times on the context stream: a launch of known
 * traffic for calibrating hardware byte counters. */
int pmp_debug_copy(pmp_ctx* ctx, int64_t bytes, int32_t repeats);

#ifdef __cplusplus
}
#endif
#endif
