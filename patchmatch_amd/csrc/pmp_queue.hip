// pmp_queue.hip -- the work-queue kernels and the batched form.
//
// Every stage is one launch of a persistent kernel whose blocks walk a list
// of work items over all slots of a descriptor set:
//   k_init_q       random init + initial score; items (slot, 64x4 tile), the
//                  tile's reference taps staged in LDS
//   k_score_q      score of the working planes; items (slot, 64x4 tile)
//   k_propagate_q  one red-black pass; items (slot, R x 64 same-colour tile),
//                  split into 8 contiguous bands, one per XCD, each taken
//                  through an atomic head (a work queue); a block drains its
//                  own XCD's band first, then the others
//   k_copyback_q, k_median_q (x2), k_postfilter_q   finalize; items (slot, row)
// Descriptors (SlotDesc) and work lists live in device memory and are read
// with scalar loads (constant address space); launches carry only pointers,
// so a captured hipGraph of a whole densification replays for any slots with
// the same iteration count.  The view count is a run-time value (views in
// groups of 3, 2, 1, per-view scores parked in LDS): one kernel serves
// keyframes of every size and view count in the same launch.
#include <math.h>
#include <string.h>

#include "pmp_trace.h"
#include "pmp_internal.h"
#include "pmp_math.h"
#include "pmp_queue.h"
#include "pmp_score.h"
#include "pmp_tile.h"

#define PMP_CONST __attribute__((address_space(4)))

namespace pmp {

// Slot holding `item` (uniform binary search) and the item's index in it.
__device__ __forceinline__ int list_find(const WorkList PMP_CONST* L, int item, int& local) {
    int lo = 0, hi = L->n - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (item < L->end[mid]) hi = mid;
        else lo = mid + 1;
    }
    local = item - (lo ? L->end[lo - 1] : 0);
    return L->slot[lo];
}

__device__ __forceinline__ const SlotDesc& slot_ref(const SlotDesc PMP_CONST* slots, int s) {
    return *(const SlotDesc*)(slots + s);
}

__global__ void k_fill_u32(unsigned int* __restrict__ p, int n, unsigned int v) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

// An init tile is 4 rows x 64 columns, one row per wave.  Its reference
// patch and bearing-table entries (the tile plus a halo of half each side,
// coordinates clamped: only interior pixels, whose taps lie inside the image,
// read them) are staged in LDS, as for a propagation tile, so that the tap
// reads never wait behind the neighbour gathers issued before them.
constexpr int kInitRows = 4;
constexpr int kInitStageRows = kInitRows + 2 * kMaxHalf;
struct InitStage {
    float2 row[kInitStageRows];
    float2 col[kTileCols];
    uint8_t ref[kInitStageRows * kTileCols];
};
static_assert(64 + 2 * kMaxHalf <= kTileCols, "init tile columns");

// Waits for the previous tile's readers first.
__device__ __forceinline__ LdsTaps stage_init_tile(const KfArgs& a, InitStage& st, int tx, int ty) {
    const int r0 = ty * kInitRows - a.half, c0 = 64 * tx - a.half;
    const int nr = kInitRows + 2 * a.half, nc = 64 + 2 * a.half;
    __syncthreads();
    for (int e = threadIdx.x; e < nc; e += 256) st.col[e] = a.coltab[min(max(c0 + e, 0), a.W - 1)];
    for (int e = threadIdx.x; e < nr; e += 256) st.row[e] = a.rowtab[min(max(r0 + e, 0), a.H - 1)];
    for (int e = threadIdx.x; e < nr * nc; e += 256) {
        const int r = e / nc, c = e - r * nc;
        const int y = min(max(r0 + r, 0), a.H - 1), x = min(max(c0 + c, 0), a.W - 1);
        st.ref[r * kTileCols + c] = a.ref[(size_t)y * a.W + x];
    }
    __syncthreads();
    return LdsTaps{st.row, st.col, st.ref, r0, c0};
}

// kScore: only score the working planes (pmp_score_plane_depth); otherwise
// random initialisation + score + snapshot.
template <bool kZero, bool kScore>
__global__ __launch_bounds__(256) void k_init_q(const SlotDesc PMP_CONST* slots, const BatchHdr PMP_CONST* hdr) {
    __shared__ float sWc[256];
    __shared__ float sv[kMaxViews][256];
    __shared__ InitStage st;
    const WorkList PMP_CONST* L = &hdr->init;
    const int total = L->total;
    int cur = -1;
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        int local;
        const int s = list_find(L, item, local);
        const SlotDesc& d = slot_ref(slots, s);
        if (s != cur) {
            __syncthreads();
            load_wc(d.a.Wc, sWc);
            cur = s;
        }
        const int tx = local % d.nbx_init, ty = local / d.nbx_init;
        const LdsTaps tp = stage_init_tile(d.a, st, tx, ty);
        const int j = tx * 64 + (threadIdx.x & 63);
        const int i = ty * kInitRows + (threadIdx.x >> 6);
        if (i >= d.a.H || j >= d.a.W) continue;
        if constexpr (kScore) {
            if (!in_score_grid(d.a, i, j)) continue;
            const size_t k = (size_t)i * d.a.W + j;
            const RefStat rs = ref_stats(d.a, sWc, i, j);
            const float4 p = d.plane[k];
            d.score[k] = score_plane_t<0, kZero>(d.a, tp, sWc, rs, i, j, p.x, p.y, p.z, &sv[0][threadIdx.x], 256);
        } else {
            init_pixel<0, kZero>(d.a, tp, sWc, i, j, d.plane, d.depth, d.score, d.iplane, d.idepth, d.iscore,
                                 &sv[0][threadIdx.x], 256);
        }
    }
}

// Debug pass epochs (SlotDesc::epochs).  A tile about to run red-black pass
// pp must have completed pp passes, and each of the 8 tiles around it (whose
// maps its candidates and the other colour's pixels come from, 3 pixels at
// most across a tile edge) pp or pp + 1: the previous pass wrote them and this
// one writes only the colour that pass pp does not read.  Anything else (a
// tile skipped, run twice, or run before the pass before it finished) adds to
// the violation count.  Tiles of one launch may see each other's stores or not;
// both values are allowed.
__device__ __forceinline__ void epoch_enter(const SlotDesc& d, int tx, int ty, int pp) {
    if (threadIdx.x != 0) return;
    unsigned int* e = d.epochs;
    const int nbx = d.nbx_prop, nby = d.nby_prop;
    unsigned int bad = 0;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const int x = tx + dx, y = ty + dy;
            if (x < 0 || y < 0 || x >= nbx || y >= nby) continue;
            const unsigned int v = __hip_atomic_load(&e[y * nbx + x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const bool ok = (dx == 0 && dy == 0) ? v == (unsigned int)pp : (v == (unsigned int)pp || v == (unsigned int)pp + 1u);
            bad += ok ? 0u : 1u;
        }
    if (bad) atomicAdd(&e[nbx * nby], bad);
}

// After every lane's map writes of the tile: its epoch becomes pp + 1.
__device__ __forceinline__ void epoch_leave(const SlotDesc& d, int tx, int ty, int pp) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        __hip_atomic_store(&d.epochs[ty * d.nbx_prop + tx], (unsigned int)pp + 1u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The zero-edge variant is held to 4 waves per SIMD (<= 128 VGPRs; left
// alone it takes 132 and runs at 3 waves, 28 % slower); the wrap/clamp
// variant fits by itself and is faster with the compiler's own allocation.
template <bool kZero, int R>
__global__ __launch_bounds__(64 * R) __attribute__((amdgpu_waves_per_eu(kZero ? 4 : 1))) void k_propagate_q(const SlotDesc PMP_CONST* slots,
                                                        const BatchHdr PMP_CONST* hdr, int pass,
                                                        unsigned int* __restrict__ heads,
                                                        unsigned long long* __restrict__ counters, int bands) {
    __shared__ PropShared<R, kMaxViews> sh;
    const WorkList PMP_CONST* L = &hdr->prop;
    const int total = L->total;
    const int nb = bands ? 8 : 1;
    const int home = bands ? (int)(blockIdx.x & 7) : 0;
    unsigned int n_eval = 0, n_cand = 0;
    int cur = -1;
    for (int k = 0; k < nb; ++k) {
        const int band = (home + k) & (nb - 1);
        const int lo = (int)(((long long)total * band) / nb);
        const int hi = (int)(((long long)total * (band + 1)) / nb);
        for (;;) {
            // every block takes items until its band's head passes the end:
            // each atomicAdd moves the head, so the loop ends for every block
            if (threadIdx.x == 0) sh.item = lo + (int)atomicAdd(&heads[band], 1u);
            __syncthreads();
            const int item = sh.item;
            __syncthreads();
            if (item >= hi) break;
            int local;
            const int s = list_find(L, item, local);
            const SlotDesc& d = slot_ref(slots, s);
            if (s != cur) {
                load_wc(d.a.Wc, sh.sWc);
                cur = s;
            }
            const int pp = d.pass0 + pass;
            const int tx = local % d.nbx_prop, ty = local / d.nbx_prop;
            if (d.epochs) epoch_enter(d, tx, ty, pp);
            propagate_tile<0, kZero, R>(d.a, sh, d.plane, d.depth, d.score, pp >> 1, pp & 1, tx, ty,
                                        64 * (item & (R - 1)), n_eval, n_cand);
            if (d.epochs) epoch_leave(d, tx, ty, pp);
        }
    }
    count_add(&counters[0], n_eval);
    count_add(&counters[1], n_cand);
}

__global__ __launch_bounds__(256) void k_copyback_q(const SlotDesc PMP_CONST* slots, const BatchHdr PMP_CONST* hdr) {
    const WorkList PMP_CONST* L = &hdr->rows;
    for (int item = blockIdx.x; item < L->total; item += gridDim.x) {
        int i;
        const SlotDesc& d = slot_ref(slots, list_find(L, item, i));
        for (int j = threadIdx.x; j < d.a.W; j += blockDim.x)
            copyback_pixel(d.a, d.mask, d.iplane, d.idepth, d.iscore, d.plane, d.depth, d.score, d.oplane, d.odepth,
                           d.oscore, d.onghbr, d.min_var, d.inv_T, i, j);
    }
}

__global__ __launch_bounds__(256) void k_median_q(const SlotDesc PMP_CONST* slots, const BatchHdr PMP_CONST* hdr,
                                                  int second) {
    const WorkList PMP_CONST* L = &hdr->rows;
    for (int item = blockIdx.x; item < L->total; item += gridDim.x) {
        int i;
        const SlotDesc& d = slot_ref(slots, list_find(L, item, i));
        if (!d.postprocess) continue;
        const float* src = second ? d.tmp0 : d.odepth;
        float* dst = second ? d.tmp1 : d.tmp0;
        for (int j = threadIdx.x; j < d.a.W; j += blockDim.x) median3_pixel(src, dst, d.a.W, d.a.H, i, j);
    }
}

__global__ __launch_bounds__(256) void k_postfilter_q(const SlotDesc PMP_CONST* slots,
                                                      const BatchHdr PMP_CONST* hdr) {
    const WorkList PMP_CONST* L = &hdr->rows;
    for (int item = blockIdx.x; item < L->total; item += gridDim.x) {
        int i;
        const SlotDesc& d = slot_ref(slots, list_find(L, item, i));
        if (!d.postprocess) continue;
        for (int j = threadIdx.x; j < d.a.W; j += blockDim.x)
            postfilter_pixel(d.tmp1, d.odepth, d.oscore, (size_t)i * d.a.W + j, d.rel_thr, d.a.min_depth,
                             d.a.max_depth, d.min_score);
    }
}

// ---------------------------------------------------------------- host side

int desc_alloc(pmp_ctx* ctx, DescSet& d, int cap) {
    d.cap = cap;
    PMP_HIP(ctx, hipHostMalloc((void**)&d.h_slots, cap * sizeof(SlotDesc), hipHostMallocDefault));
    PMP_HIP(ctx, hipHostMalloc((void**)&d.h_hdr, sizeof(BatchHdr), hipHostMallocDefault));
    PMP_HIP(ctx, hipMalloc(&d.d_slots, cap * sizeof(SlotDesc)));
    PMP_HIP(ctx, hipMalloc(&d.d_hdr, sizeof(BatchHdr)));
    PMP_HIP(ctx, hipMalloc(&d.d_heads, 8 * kMaxPasses * sizeof(unsigned int)));
    PMP_HIP(ctx, hipEventCreateWithFlags(&d.uploaded, hipEventDisableTiming));
    return PMP_OK;
}

void desc_free(DescSet& d) {
    for (auto& kv : d.graphs) {
        (void)hipGraphExecDestroy(kv.second.second);
        (void)hipGraphDestroy(kv.second.first);
    }
    if (d.h_slots) (void)hipHostFree(d.h_slots);
    if (d.h_hdr) (void)hipHostFree(d.h_hdr);
    if (d.d_slots) (void)hipFree(d.d_slots);
    if (d.d_hdr) (void)hipFree(d.d_hdr);
    if (d.d_heads) (void)hipFree(d.d_heads);
    if (d.uploaded) (void)hipEventDestroy(d.uploaded);
    d = DescSet();
}

int desc_begin(pmp_ctx* ctx, DescSet& d) {
    if (d.pending) {
        PMP_HIP(ctx, hipEventSynchronize(d.uploaded));
        d.pending = false;
    }
    return PMP_OK;
}

int desc_upload(pmp_ctx* ctx, DescSet& d, int n, hipStream_t stream) {
    PMP_HIP(ctx, hipMemcpyAsync(d.d_slots, d.h_slots, n * sizeof(SlotDesc), hipMemcpyHostToDevice, stream));
    PMP_HIP(ctx, hipMemcpyAsync(d.d_hdr, d.h_hdr, sizeof(BatchHdr), hipMemcpyHostToDevice, stream));
    PMP_HIP(ctx, hipEventRecord(d.uploaded, stream));
    d.pending = true;
    return PMP_OK;
}

void queue_free(pmp_ctx* ctx) {
    QueueState* q = ctx->queue;
    if (!q) return;
    desc_free(q->plain);
    desc_free(q->batch);
    for (auto& b : q->bufs) free_slot(b);
    delete q;
    ctx->queue = nullptr;
}

int queue_state(pmp_ctx* ctx, QueueState*& out) {
    if (!ctx->queue) {
        QueueState* q = new (std::nothrow) QueueState();
        if (!q) return set_error(ctx, PMP_ENOMEM, "queue state");
        ctx->queue = q;
        int rc = desc_alloc(ctx, q->plain, 1);
        if (!rc) rc = desc_alloc(ctx, q->batch, kMaxSlots);
        if (rc) {
            queue_free(ctx);
            return rc;
        }
        PMP_HIP(ctx, hipDeviceGetAttribute(&q->cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
        q->bufs.resize(kMaxSlots);
    }
    out = ctx->queue;
    return PMP_OK;
}

void fill_list(WorkList& L, int n, const int* items) {
    L.n = 0;
    L.total = 0;
    for (int s = 0; s < n; ++s) {
        if (items[s] <= 0) continue;
        L.total += items[s];
        L.slot[L.n] = s;
        L.end[L.n] = L.total;
        L.n++;
    }
}

int slot_fill(pmp_ctx* ctx, const pmp_params* params, const pmp_keyframe& kf, const SlotBuf& b, const uint8_t* dmask,
              int pass0, int rows, hipStream_t stream, SlotDesc& d, int items[3]) {
    pmp_params pk = *params;
    if (kf.min_depth > 0.0f) pk.min_depth = kf.min_depth;
    if (kf.max_depth > 0.0f) pk.max_depth = kf.max_depth;
    memset(&d, 0, sizeof(d));
    int rc = build_args(ctx, &pk, kf.ref_id, kf.n_views, kf.view_ids, (const double*)&kf.ref_pose,
                        (const double*)kf.view_poses, stream, d.a);
    if (rc) return rc;
    const int W = d.a.W, H = d.a.H;
    if ((size_t)W * H > b.cap) return set_error(ctx, PMP_ESTATE, "keyframe maps are smaller than the keyframe");
    d.plane = b.plane; d.depth = b.depth; d.score = b.score;
    d.iplane = b.iplane; d.idepth = b.idepth; d.iscore = b.iscore;
    d.oplane = b.oplane; d.odepth = b.odepth; d.oscore = b.oscore; d.onghbr = b.onghbr;
    d.tmp0 = b.tmp0; d.tmp1 = b.tmp1;
    d.mask = dmask;
    if (!dmask && kf.mask_id >= 0) {
        auto im = ctx->images.find((uint64_t)kf.mask_id);
        if (im == ctx->images.end()) return set_error(ctx, PMP_ENOTFOUND, "unknown mask image id");
        if (im->second.W != W || im->second.H != H)
            return set_error(ctx, PMP_EINVAL, "mask image size differs from the reference");
        d.mask = im->second.gray;
    }
    d.min_var = pk.min_patch_sd * pk.min_patch_sd;
    d.inv_T = (float)(1.0 / (double)d.a.T);
    d.rel_thr = pk.median_rel_thr;
    d.min_score = pk.min_score;
    d.postprocess = pk.postprocess;
    d.pass0 = pass0;
    d.nbx_init = (W + 63) / 64;
    d.nbx_prop = ((W + 1) / 2 + kTileSame - 1) / kTileSame;
    items[0] = d.nbx_init * ((H + 3) / 4);
    const int trows = 64 * rows / kTileSame;   // rows of a propagation tile
    d.nby_prop = (H + trows - 1) / trows;
    items[1] = d.nbx_prop * d.nby_prop;
    items[2] = H;
    return PMP_OK;
}

static const void* prop_kernel(bool zero, int rows) {
    if (zero) {
        if (rows == 2) return (const void*)k_propagate_q<true, 2>;
        if (rows == 4) return (const void*)k_propagate_q<true, 4>;
        return (const void*)k_propagate_q<true, 8>;
    }
    if (rows == 2) return (const void*)k_propagate_q<false, 2>;
    if (rows == 4) return (const void*)k_propagate_q<false, 4>;
    return (const void*)k_propagate_q<false, 8>;
}

static int check_config(pmp_ctx* ctx, const pmp_batch_config& c) {
    if (c.tile_rows != 2 && c.tile_rows != 4 && c.tile_rows != 8)
        return set_error(ctx, PMP_EINVAL, "tile_rows must be 2, 4 or 8");
    if (c.blocks_per_cu < 0 || c.blocks_per_cu > 32) return set_error(ctx, PMP_EINVAL, "blocks_per_cu must be in [0, 32]");
    return PMP_OK;
}

// Persistent grid: the configured blocks per CU, capped by (and defaulting
// to) what fits on a CU at once; a multiple of 8 with per-XCD bands.
int queue_run(pmp_ctx* ctx, const QueueState& q, const pmp_batch_config& c, bool zero, bool post, QueueRun& r) {
    int occ = 0;
    PMP_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, prop_kernel(zero, c.tile_rows),
                                                              64 * c.tile_rows, 0));
    if (occ < 1) occ = 1;
    const int bpc = (c.blocks_per_cu <= 0 || c.blocks_per_cu > occ) ? occ : c.blocks_per_cu;
    r.zero = zero;
    r.post = post;
    r.rows = c.tile_rows;
    r.bands = c.xcd_bands ? 1 : 0;
    r.blocks = q.cus * bpc;
    if (r.bands) r.blocks = (r.blocks + 7) / 8 * 8;
    return PMP_OK;
}

#define PMP_QPTRS                                                             \
    const SlotDesc PMP_CONST* sl = (const SlotDesc PMP_CONST*)d.d_slots;     \
    const BatchHdr PMP_CONST* hd = (const BatchHdr PMP_CONST*)d.d_hdr;

static int grid_blocks(pmp_ctx* ctx) {
    QueueState* q = ctx->queue;
    return (q && q->cus > 0 ? q->cus : 256) * 8;
}

int q_clear(pmp_ctx* ctx, hipStream_t s, DescSet& d, int passes) {
    // fill kernels, not memset nodes: the sequence may be captured
    const int n = 8 * passes;
    if (n > 0)
        hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d.d_heads, n, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, (unsigned int*)ctx->counters, 4, 0u);
    PMP_HIP(ctx, hipGetLastError());
    return PMP_OK;
}

int q_init(pmp_ctx* ctx, hipStream_t s, DescSet& d, const QueueRun& r) {
    PMP_QPTRS
    const dim3 grd((unsigned)grid_blocks(ctx)), blk(256);
    if (r.zero) hipLaunchKernelGGL((k_init_q<true, false>), grd, blk, 0, s, sl, hd);
    else hipLaunchKernelGGL((k_init_q<false, false>), grd, blk, 0, s, sl, hd);
    PMP_HIP(ctx, hipGetLastError());
    return PMP_OK;
}

int q_score(pmp_ctx* ctx, hipStream_t s, DescSet& d, const QueueRun& r) {
    PMP_QPTRS
    const dim3 grd((unsigned)grid_blocks(ctx)), blk(256);
    if (r.zero) hipLaunchKernelGGL((k_init_q<true, true>), grd, blk, 0, s, sl, hd);
    else hipLaunchKernelGGL((k_init_q<false, true>), grd, blk, 0, s, sl, hd);
    PMP_HIP(ctx, hipGetLastError());
    return PMP_OK;
}

int q_prop(pmp_ctx* ctx, hipStream_t s, DescSet& d, const QueueRun& r, int pass) {
    PMP_QPTRS
    const dim3 grd((unsigned)r.blocks), blk((unsigned)(64 * r.rows));
    unsigned int* heads = d.d_heads + 8 * pass;
#define PMP_PROP(Z, R) \
    hipLaunchKernelGGL((k_propagate_q<Z, R>), grd, blk, 0, s, sl, hd, pass, heads, ctx->counters, r.bands)
    if (r.zero) {
        if (r.rows == 2) PMP_PROP(true, 2);
        else if (r.rows == 4) PMP_PROP(true, 4);
        else PMP_PROP(true, 8);
    } else {
        if (r.rows == 2) PMP_PROP(false, 2);
        else if (r.rows == 4) PMP_PROP(false, 4);
        else PMP_PROP(false, 8);
    }
#undef PMP_PROP
    PMP_HIP(ctx, hipGetLastError());
    return PMP_OK;
}

int q_finalize(pmp_ctx* ctx, hipStream_t s, DescSet& d, const QueueRun& r) {
    PMP_QPTRS
    const dim3 grd((unsigned)grid_blocks(ctx)), blk(256);
    hipLaunchKernelGGL(k_copyback_q, grd, blk, 0, s, sl, hd);
    if (r.post) {
        hipLaunchKernelGGL(k_median_q, grd, blk, 0, s, sl, hd, 0);
        hipLaunchKernelGGL(k_median_q, grd, blk, 0, s, sl, hd, 1);
        hipLaunchKernelGGL(k_postfilter_q, grd, blk, 0, s, sl, hd);
    }
    PMP_HIP(ctx, hipGetLastError());
    return PMP_OK;
}

int q_fill_u32(pmp_ctx* ctx, hipStream_t s, unsigned int* p, size_t n, unsigned int v) {
    for (size_t off = 0; off < n; off += (size_t)1 << 30) {
        const size_t m = (n - off) < ((size_t)1 << 30) ? (n - off) : ((size_t)1 << 30);
        hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, p + off, (int)m, v);
    }
    PMP_HIP(ctx, hipGetLastError());
    return PMP_OK;
}

int epochs_arm(pmp_ctx* ctx, hipStream_t s, int n, SlotDesc* slots, const int* tiles, int pass0) {
    if (!ctx->epoch_mode) return PMP_OK;
    size_t need = 0;
    for (int k = 0; k < n; ++k) need += (size_t)tiles[k] + 1;
    if (need > ctx->epochs_cap) {
        if (ctx->epochs) (void)hipFree(ctx->epochs);
        ctx->epochs = nullptr;
        ctx->epochs_cap = 0;
        PMP_HIP(ctx, hipMalloc(&ctx->epochs, need * sizeof(unsigned int)));
        ctx->epochs_cap = need;
    }
    // fill kernels on the stream, ahead of the (possibly graph-replayed) passes
    size_t off = 0;
    int rc;
    for (int k = 0; k < n; ++k) {
        unsigned int* seg = ctx->epochs + off;
        if ((rc = q_fill_u32(ctx, s, seg, (size_t)tiles[k], (unsigned int)pass0))) return rc;
        if ((rc = q_fill_u32(ctx, s, seg + tiles[k], 1, 0u))) return rc;
        if (k == 0 && ctx->epoch_mode == 2 && (rc = q_fill_u32(ctx, s, seg, 1, (unsigned int)pass0 + 1u))) return rc;
        slots[k].epochs = seg;
        off += (size_t)tiles[k] + 1;
    }
    ctx->epoch_tiles.assign(tiles, tiles + n);
    ctx->epoch_words = need;
    return PMP_OK;
}

static int q_sequence(pmp_ctx* ctx, hipStream_t s, DescSet& d, const QueueRun& r, int passes) {
    int rc;
    if ((rc = q_clear(ctx, s, d, passes))) return rc;
    int tp = timing_begin(ctx, kStageInit);
    if ((rc = q_init(ctx, s, d, r))) return rc;
    timing_end(ctx, tp);
    for (int p = 0; p < passes; ++p) {
        tp = timing_begin(ctx, kStagePropagate);
        if ((rc = q_prop(ctx, s, d, r, p))) return rc;
        timing_end(ctx, tp);
    }
    tp = timing_begin(ctx, kStageFinalize);
    rc = q_finalize(ctx, s, d, r);
    timing_end(ctx, tp);
    return rc;
}

int q_densify(pmp_ctx* ctx, hipStream_t s, DescSet& d, const QueueRun& r, int passes, bool use_graph) {
    // per-stage timing needs events between the launches: no graph then
    if (!use_graph || ctx->timing) return q_sequence(ctx, s, d, r, passes);
    const GraphKey key{passes, r.zero ? 1 : 0, r.rows, r.blocks, r.bands, r.post ? 1 : 0};
    auto it = d.graphs.find(key);
    if (it == d.graphs.end()) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        PMP_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = q_sequence(ctx, s, d, r, passes);
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return check_hip(ctx, e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            return check_hip(ctx, e, "hipGraphInstantiate");
        }
        it = d.graphs.emplace(key, std::make_pair(g, ge)).first;
    }
    PMP_HIP(ctx, hipGraphLaunch(it->second.second, s));
    return PMP_OK;
}

// ---------------------------------------------------------------- batched form

// Descriptors and work lists for n keyframes (all from pass 0), uploaded to
// the batch set.
// arm: pass epochs for this batch (pmp_batch_densify; the autotuner reruns
// passes from the same maps and does not arm them)
static int batch_prepare(pmp_ctx* ctx, QueueState* q, const pmp_params* params, int n, const pmp_keyframe* kfs,
                         int rows, bool arm) {
    int rc = check_params(ctx, params);
    if (rc) return rc;
    if (params->iterations > kMaxPasses / 2) return set_error(ctx, PMP_EINVAL, "iterations must be <= 64");
    // the slots' maps may be reallocated and the staging copy rewritten
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if ((rc = desc_begin(ctx, q->batch))) return rc;
    if (!ctx->counters) {
        PMP_HIP(ctx, hipMalloc(&ctx->counters, 4 * sizeof(unsigned long long)));
        PMP_HIP(ctx, hipMemsetAsync(ctx->counters, 0, 4 * sizeof(unsigned long long), ctx->stream));
    }
    int it_init[kMaxSlots], it_prop[kMaxSlots], it_rows[kMaxSlots];
    for (int k = 0; k < n; ++k) {
        auto ir = ctx->images.find(kfs[k].ref_id);
        if (ir == ctx->images.end()) return set_error(ctx, PMP_ENOTFOUND, "unknown reference image id");
        if ((rc = ensure_slot(ctx, q->bufs[k], ir->second.W, ir->second.H))) return rc;
        int items[3];
        rc = slot_fill(ctx, params, kfs[k], q->bufs[k], nullptr, 0, rows, ctx->stream, q->batch.h_slots[k], items);
        if (rc) return rc;
        it_init[k] = items[0];
        it_prop[k] = items[1];
        it_rows[k] = items[2];
    }
    fill_list(q->batch.h_hdr->init, n, it_init);
    fill_list(q->batch.h_hdr->prop, n, it_prop);
    fill_list(q->batch.h_hdr->rows, n, it_rows);
    if (arm && (rc = epochs_arm(ctx, ctx->stream, n, q->batch.h_slots, it_prop, 0))) return rc;
    return desc_upload(ctx, q->batch, n, ctx->stream);
}

}  // namespace pmp

using namespace pmp;

extern "C" {

int pmp_batch_densify(pmp_ctx* ctx, const pmp_params* params, int32_t n, const pmp_keyframe* kfs) {
    TraceRange trace_range("pmp_batch_densify");
    if (!ctx || !params || !kfs) return PMP_EINVAL;
    if (n < 1 || n > kMaxSlots) return set_error(ctx, PMP_EINVAL, "batch size must be in [1, PMP_MAX_BATCH]");
    PMP_ON_DEVICE(ctx);
    QueueState* q = nullptr;
    int rc = queue_state(ctx, q);
    if (rc) return rc;
    q->n_last = 0;
    if ((rc = batch_prepare(ctx, q, params, n, kfs, q->cfg.tile_rows, true))) return rc;
    QueueRun r;
    if ((rc = queue_run(ctx, *q, q->cfg, params->edge_mode == PMP_EDGE_ZERO, params->postprocess != 0, r))) return rc;
    if ((rc = q_densify(ctx, ctx->stream, q->batch, r, 2 * params->iterations, q->cfg.use_graph != 0))) return rc;
    q->n_last = n;
    ctx->densify_calls += n;
    return PMP_OK;
}

int pmp_batch_get_maps(pmp_ctx* ctx, int32_t k, float* plane4, float* depth, float* score, int32_t* nghbr) {
    if (!ctx) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    QueueState* q = ctx->queue;
    if (!q || k < 0 || k >= q->n_last) return set_error(ctx, PMP_EINVAL, "no such keyframe in the last batch");
    const SlotBuf& b = q->bufs[k];
    const size_t P = (size_t)b.W * b.H;
    if (plane4) PMP_HIP(ctx, hipMemcpyAsync(plane4, b.oplane, P * 16, hipMemcpyDeviceToHost, ctx->stream));
    if (depth) PMP_HIP(ctx, hipMemcpyAsync(depth, b.odepth, P * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (score) PMP_HIP(ctx, hipMemcpyAsync(score, b.oscore, P * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (nghbr) PMP_HIP(ctx, hipMemcpyAsync(nghbr, b.onghbr, P * 4, hipMemcpyDeviceToHost, ctx->stream));
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PMP_OK;
}

int pmp_batch_set_config(pmp_ctx* ctx, const pmp_batch_config* cfg) {
    if (!ctx || !cfg) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    int rc = check_config(ctx, *cfg);
    if (rc) return rc;
    QueueState* q = nullptr;
    if ((rc = queue_state(ctx, q))) return rc;
    q->cfg = *cfg;
    return PMP_OK;
}

int pmp_batch_get_config(pmp_ctx* ctx, pmp_batch_config* cfg) {
    if (!ctx || !cfg) return PMP_EINVAL;
    PMP_ON_DEVICE(ctx);
    QueueState* q = nullptr;
    int rc = queue_state(ctx, q);
    if (rc) return rc;
    *cfg = q->cfg;
    return PMP_OK;
}

int pmp_batch_autotune(pmp_ctx* ctx, const pmp_params* params, int32_t n, const pmp_keyframe* kfs,
                       pmp_batch_config* best) {
    TraceRange trace_range("pmp_batch_autotune");
    if (!ctx || !params || !kfs) return PMP_EINVAL;
    if (n < 1 || n > kMaxSlots) return set_error(ctx, PMP_EINVAL, "batch size must be in [1, PMP_MAX_BATCH]");
    PMP_ON_DEVICE(ctx);
    QueueState* q = nullptr;
    int rc = queue_state(ctx, q);
    if (rc) return rc;
    q->n_last = 0;
    hipEvent_t e0, e1;
    PMP_HIP(ctx, hipEventCreate(&e0));
    PMP_HIP(ctx, hipEventCreate(&e1));
    pmp_batch_config found = q->cfg;
    float best_ms = 3.0e38f;
    const int rows_set[3] = {2, 4, 8};
    const bool zero = params->edge_mode == PMP_EDGE_ZERO;
    for (int ri = 0; ri < 3 && rc == 0; ++ri) {
        if ((rc = batch_prepare(ctx, q, params, n, kfs, rows_set[ri], false))) break;
        pmp_batch_config c = q->cfg;
        c.tile_rows = rows_set[ri];
        c.blocks_per_cu = 32;
        QueueRun r;
        if ((rc = queue_run(ctx, *q, c, zero, false, r))) break;
        const int occ = r.blocks / (q->cus > 0 ? q->cus : 1);   // blocks per CU that fit (rounding aside)
        // at least 8 waves per CU (rows waves per block): fewer leave SIMDs
        // idle and measured 1.3-4x slower, and their trials are the longest
        const int bpc_min = std::min(occ, (8 + rows_set[ri] - 1) / rows_set[ri]);
        for (int bpc = bpc_min; bpc <= occ && rc == 0; ++bpc) {
            for (int bands = 0; bands <= 1 && rc == 0; ++bands) {
                c.blocks_per_cu = bpc;
                c.xcd_bands = bands;
                if ((rc = queue_run(ctx, *q, c, zero, false, r))) break;
                // same starting maps for every candidate: init, then time every
                // pass.  The first passes score random planes (most candidates
                // survive pre-rejection, the gathers scatter); the converged
                // ones pre-reject most candidates and are instruction-bound.
                // The shapes rank differently on the two kinds, so a prefix of
                // passes misranks them.
                const int timed = 2 * params->iterations;
                if ((rc = q_clear(ctx, ctx->stream, q->batch, timed))) break;   // every timed pass's heads
                if ((rc = q_init(ctx, ctx->stream, q->batch, r))) break;
                hipError_t e = hipEventRecord(e0, ctx->stream);
                for (int pass = 0; pass < timed && e == hipSuccess && rc == 0; ++pass)
                    rc = q_prop(ctx, ctx->stream, q->batch, r, pass);
                if (e == hipSuccess && rc == 0) e = hipEventRecord(e1, ctx->stream);
                if (e == hipSuccess && rc == 0) e = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (e == hipSuccess && rc == 0) e = hipEventElapsedTime(&ms, e0, e1);
                if (e != hipSuccess) rc = check_hip(ctx, e, "autotune timing");
                if (rc) break;
                if (ms < best_ms) {
                    best_ms = ms;
                    found = c;
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    q->cfg = found;
    if (best) *best = found;
    return PMP_OK;
}

}  // extern "C"
