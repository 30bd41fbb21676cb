// pmp_estimator.hip -- keyframe densification kernels and their C-ABI.
//
// Kernels (one launch each unless noted):
//   k_pack        uint8 image -> 2x2-neighbourhood words (one 32-bit gather per
//                 bilinear sample)                                  HBM-bound
//   k_init        random hypothesis + initial score + snapshot   VALU-bound
//                 (PatchMatchCuda.cu:210-270)
//   k_score       score of the working plane map (PatchMatchCuda.cu:255-270)
//   k_propagate   one red-black pass (PatchMatchCuda.cu:273-407), 2 per
//                 iteration                                        VALU-bound
//   k_copyback    copy-back window / gate (depthmap.cc:673-687, :731-760)
//   k_median3     3x3 median, replicate border (cv::medianBlur), 2 launches
//   k_postfilter  PostProcess test (depthmap.cc:1176-1198)            HBM-bound
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdio>
#include <string>

#include "pmp_internal.h"
#include "pmp_math.h"
#include "pmp_score.h"

namespace pmp {

__device__ __forceinline__ void load_wc(const KfArgs& a, float* sWc) {
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    for (int k = tid; k < 256; k += blockDim.x * blockDim.y) sWc[k] = a.Wc[k];
    __syncthreads();
}

__device__ __forceinline__ void count_add(unsigned long long* ctr, unsigned int v) {
    // one atomic per wave (all 64 lanes active; per-lane counts are < 2^16)
    unsigned int s = v;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(ctr, (unsigned long long)s);
}

// ---------------------------------------------------------------- pack
__global__ void k_pack(const uint8_t* __restrict__ g, uint32_t* __restrict__ q, int W, int H, int wrap) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W || y >= H) return;
    int x1 = x + 1 < W ? x + 1 : (wrap ? 0 : W - 1);
    int y1 = y + 1 < H ? y + 1 : H - 1;
    const uint8_t* r0 = g + (size_t)y * W;
    const uint8_t* r1 = g + (size_t)y1 * W;
    q[(size_t)y * W + x] = (uint32_t)r0[x] | ((uint32_t)r0[x1] << 8) | ((uint32_t)r1[x] << 16) |
                           ((uint32_t)r1[x1] << 24);
}

// ---------------------------------------------------------------- tiles
// Workgroups are dispatched to the 8 XCDs round-robin by linear block id.  The
// grid is padded to a multiple of 8 and block b is given tile
// (b mod 8) * (nb8/8) + b/8, so each XCD walks one contiguous band of tiles
// and its L2 holds that band's neighbour rows.  Returns -1 for padding blocks.
__device__ __forceinline__ int xcd_tile(int b, int nb8, int nb) {
    const int t = (b & 7) * (nb8 >> 3) + (b >> 3);
    return t < nb ? t : -1;
}

// ---------------------------------------------------------------- init
// Tiles of 64 x 4 pixels, one thread per pixel.
template <int NV, bool kZero>
__global__ __launch_bounds__(256) void k_init(KfArgs a, float4* __restrict__ plane, float* __restrict__ depth,
                                              float* __restrict__ score, float4* __restrict__ iplane,
                                              float* __restrict__ idepth, float* __restrict__ iscore, int nbx,
                                              int nb) {
    __shared__ float sWc[256];
    const int tile = xcd_tile(blockIdx.x, gridDim.x, nb);
    if (tile < 0) return;
    load_wc(a, sWc);
    const int j = (tile % nbx) * 64 + (threadIdx.x & 63);
    const int i = (tile / nbx) * 4 + (threadIdx.x >> 6);
    if (i >= a.H || j >= a.W) return;
    float p0 = 0.0f, p1 = 0.0f, p2 = 0.0f, dep = 0.0f;
    if (in_init_grid(a, i, j)) random_plane(a, i, j, p0, p1, p2, dep);
    float sc = __uint_as_float(0xFFFFFFFFu);   // cudaMemset(score, 0xFF) (depthmap.cc:657)
    if (in_score_grid(a, i, j)) {
        RefStat rs = ref_stats(a, sWc, i, j);
        sc = score_plane<NV, kZero>(a, sWc, rs, i, j, p0, p1, p2);
    }
    const size_t k = (size_t)i * a.W + j;
    const float4 pl = make_float4(p0, p1, p2, 0.0f);
    plane[k] = pl; depth[k] = dep; score[k] = sc;
    iplane[k] = pl; idepth[k] = dep; iscore[k] = sc;
}

// ---------------------------------------------------------------- score field
template <int NV, bool kZero>
__global__ __launch_bounds__(256) void k_score(KfArgs a, const float4* __restrict__ plane,
                                               float* __restrict__ score, int nbx, int nb) {
    __shared__ float sWc[256];
    const int tile = xcd_tile(blockIdx.x, gridDim.x, nb);
    if (tile < 0) return;
    load_wc(a, sWc);
    const int j = (tile % nbx) * 64 + (threadIdx.x & 63);
    const int i = (tile / nbx) * 4 + (threadIdx.x >> 6);
    if (i >= a.H || j >= a.W || !in_score_grid(a, i, j)) return;
    const size_t k = (size_t)i * a.W + j;
    RefStat rs = ref_stats(a, sWc, i, j);
    const float4 p = plane[k];
    score[k] = score_plane<NV, kZero>(a, sWc, rs, i, j, p.x, p.y, p.z);
}

// ---------------------------------------------------------------- red-black pass
// A tile is 4 rows x 64 same-colour pixels (128 columns); thread (w, l) of the
// 256-thread block updates row i = 4 by + w, column j = 2 (64 bx + l) + ((i +
// colour) & 1), i.e. every pixel of colour (i+j)&1 == colour once.  Candidates
// are at odd distance (1 or 3), so they have the other colour and are not
// written by this launch.
//
// Propagation candidates that survive the exact pre-rejection of a pixel are
// scored by whichever threads of the block are free: per candidate index the
// surviving (pixel, plane) jobs are packed into the first waves through LDS,
// scored there, and the raw scores handed back to the owning threads, which
// apply the penalty and the acceptance test in candidate order.  A pixel's
// sequence of decisions is unchanged; only the lanes that compute its scores
// differ, so results are identical to the one-thread-per-pixel loop.
template <int NV, bool kZero>
__global__ __launch_bounds__(256) void k_propagate(KfArgs a, float4* __restrict__ plane,
                                                   float* __restrict__ depth, float* __restrict__ score, int it,
                                                   int colour, unsigned long long* __restrict__ counters, int nbx,
                                                   int nb) {
    __shared__ float sWc[256];
    __shared__ float4 s_job[256];
    __shared__ float4 s_rs[256];
    __shared__ float s_ic[256];
    __shared__ float s_res[256];
    __shared__ int s_cnt[4];
    const int tile = xcd_tile(blockIdx.x, gridDim.x, nb);
    if (tile < 0) return;
    load_wc(a, sWc);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tbx = tile % nbx, tby = tile / nbx;
    const int i = tby * 4 + wave;
    const int j = 2 * (tbx * 64 + lane) + ((i + colour) & 1);
    const bool active = i < a.H && j < a.W && in_prop_grid(a, i, j);
    const int W = a.W;
    const size_t k = active ? (size_t)i * W + j : 0;
    unsigned int n_eval = 0, n_cand = 0;

    RefStat rs = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float sc = 0.0f, h0 = 0.0f, h1 = 0.0f, h2 = 0.0f, bx = 0.0f, by = 0.0f, bz = 0.0f, d0 = 1.0f, inv_d0 = 1.0f;
    if (active) {
        rs = ref_stats(a, sWc, i, j);
        sc = score[k];
        const float4 pk = plane[k];
        h0 = pk.x; h1 = pk.y; h2 = pk.z;
        bearing(a, i, j, bx, by, bz);
        d0 = depth[k];
        inv_d0 = rcp(d0);
    }
    s_rs[tid] = make_float4(rs.mx, rs.inv, rs.mxc, rs.vx);
    s_ic[tid] = rs.Ic;
    // jobs fill waves starting at a block-dependent wave, spreading the busy
    // waves of co-resident blocks over the SIMDs
    const int rot = 64 * (blockIdx.x & 3);

    // 1. propagation (PatchMatchCuda.cu:311-338)
    // candidate order (0,+1) (+1,0) (0,-1) (-1,0) (0,+3) (+3,0) (0,-3) (-3,0)
#pragma unroll 1
    for (int c = 0; c < 8; ++c) {
        bool need = false;
        float4 pn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float f = 0.0f;
        if (active) {
            const int dist = (c < 4) ? 1 : 3, dir = c & 3;
            const int di = (dir == 1) ? dist : (dir == 3) ? -dist : 0;
            const int dj = (dir == 0) ? dist : (dir == 2) ? -dist : 0;
            pn = plane[(size_t)(i + di) * W + (j + dj)];
            ++n_cand;
            const float s = fma_(pn.z, bz, fma_(pn.y, by, pn.x * bx));
            const float dn = rcp_pos(clampf(abs_(s), 1e-6f, 1e30f));
            const float rel = (d0 - dn) * inv_d0;
            const float fd = exp_((rel * rel) * -a.depth_penalty);
            const float prod = fma_(h2, pn.z, fma_(h1, pn.y, h0 * pn.x));
            const float ex = h0 - pn.x, ey = h1 - pn.y, ez = h2 - pn.z;
            const float nrm = sqrt_(fma_(ez, ez, fma_(ey, ey, ex * ex)));
            // acosf(prod/norm) is NaN, i.e. the candidate is rejected, unless |prod| <= norm
            if (nrm > 0.0f && abs_(prod) <= nrm) {
                const float ratio = clampf(prod * rcp_pos(clampf(nrm, 1e-30f, 1e30f)), -1.0f, 1.0f);
                const float ac = acos_(ratio);
                const float fp = exp_((ac * ac) * -a.plane_penalty);
                f = (1.0f - fd) * (1.0f - fp);
                // |score| <= 1: the penalised score cannot exceed f
                need = f > sc;
            }
        }
        // pack the jobs of this candidate index
        const unsigned long long bal = __ballot(need);
        if (lane == 0) s_cnt[wave] = __popcll(bal);
        __syncthreads();
        int off = 0, total = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const int cw = s_cnt[w];
            off += (w < wave) ? cw : 0;
            total += cw;
        }
        if (need) {
            const int slot = off + __popcll(bal & ((1ull << lane) - 1ull));
            s_job[(slot + rot) & 255] = make_float4(pn.x, pn.y, pn.z, __int_as_float(tid));
        }
        __syncthreads();
        if (((tid - rot) & 255) < total) {
            const float4 jb = s_job[tid];
            const int jl = __float_as_int(jb.w);
            const int il = tby * 4 + (jl >> 6);
            const int jc = 2 * (tbx * 64 + (jl & 63)) + ((il + colour) & 1);
            const float4 q = s_rs[jl];
            RefStat r2;
            r2.Ic = s_ic[jl]; r2.mx = q.x; r2.inv = q.y; r2.mxc = q.z; r2.vx = q.w;
            s_res[jl] = score_plane<NV, kZero>(a, sWc, r2, il, jc, jb.x, jb.y, jb.z);
        }
        __syncthreads();
        if (need) {
            const float sh = s_res[tid] * f;
            ++n_eval;
            if (sh > sc) { h0 = pn.x; h1 = pn.y; h2 = pn.z; sc = sh; }
        }
    }
    if (active) {
        // 2. refinement (PatchMatchCuda.cu:340-395)
        float dcur = depth_of_plane(bx, by, bz, h0, h1, h2);
        float c0 = h0, c1 = h1, c2 = h2;
        float z[3 * kMaxRefine];
        refine_normals<3 * kMaxRefine>(a, i, j, it, 3 * a.n_refine, z);
#pragma unroll 1
        for (int t = 0; t < a.n_refine; ++t) {
            // z is indexed by the loop counter: select from the unrolled array
            float ra = 0.0f, rb = 0.0f, rc = 0.0f;
#pragma unroll
            for (int u = 0; u < kMaxRefine; ++u)
                if (u == t) { ra = z[3 * u]; rb = z[3 * u + 1]; rc = z[3 * u + 2]; }
            const float dh = dcur * exp_(a.dr[t] * ra);
            const float nx = fma_(a.nr[t], ra, c0);
            const float ny = fma_(a.nr[t], rb, c1);
            const float nz = fma_(a.nr[t], rc, c2);
            float q0, q1, q2;
            plane_from_depth_normal(bx, by, bz, dh, nx, ny, nz, q0, q1, q2);
            const float sh = score_plane<NV, kZero>(a, sWc, rs, i, j, q0, q1, q2);
            ++n_eval;
            ++n_cand;
            if (sh > sc) { c0 = q0; c1 = q1; c2 = q2; sc = sh; dcur = dh; }
        }
        score[k] = sc;
        plane[k] = make_float4(c0, c1, c2, 0.0f);
        depth[k] = depth_of_plane(bx, by, bz, c0, c1, c2);
    }
    count_add(&counters[0], n_eval);
    count_add(&counters[1], n_cand);
}

// ---------------------------------------------------------------- finalize
// Copy-back of ComputePatchMatch: masked pixels keep the AssignMatrices values
// (depth 0, plane 0, score -1, nghbr 0); unmasked pixels take the maps after
// RandomInitialization (depthmap.cc:673-687) unless they are inside the window
// [ps-2, dim-ps-2) (depthmap.cc:731-732), where they take the final maps, or
// 0 when the patch variance is below min_patch_sd^2 (depthmap.cc:739-743).
__global__ void k_copyback(KfArgs a, const uint8_t* __restrict__ mask, const float4* __restrict__ ip,
                           const float* __restrict__ id, const float* __restrict__ is,
                           const float4* __restrict__ fp, const float* __restrict__ fd,
                           const float* __restrict__ fs, float4* __restrict__ op, float* __restrict__ od,
                           float* __restrict__ os, int32_t* __restrict__ on, float min_var, float inv_T) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= a.W) return;
    const int W = a.W, H = a.H;
    const size_t k = (size_t)i * W + j;
    const int ps = a.half + 2;
    const int lo = ps - 2, hi_r = H - ps - 2, hi_c = W - ps - 2;
    float4 pl = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float dep = 0.0f, sc = -1.0f;
    int ng = 0;
    if (mask == nullptr || mask[k]) {
        if (i >= lo && i < hi_r && j >= lo && j < hi_c) {
            bool low = false;
            if (min_var > 0.0f) {
                float sum = 0.0f;
                for (int k1 = -a.half; k1 <= a.half; k1 += a.stride)
                    for (int k2 = -a.half; k2 <= a.half; k2 += a.stride)
                        sum = sum + (float)a.ref[(size_t)(i + k1) * W + (j + k2)];
                const float mean = sum * inv_T;
                float s2 = 0.0f;
                for (int k1 = -a.half; k1 <= a.half; k1 += a.stride)
                    for (int k2 = -a.half; k2 <= a.half; k2 += a.stride) {
                        float e = (float)a.ref[(size_t)(i + k1) * W + (j + k2)] - mean;
                        s2 = fma_(e, e, s2);
                    }
                low = (s2 * inv_T) < min_var;
            }
            if (!low) {
                dep = fd[k]; sc = fs[k];
                const float4 f = fp[k];
                pl = make_float4(f.x, f.y, f.z, 0.0f);
            } else {
                sc = 0.0f;
            }
            ng = 1;
        } else {
            dep = id[k]; sc = is[k];
            const float4 f = ip[k];
            pl = make_float4(f.x, f.y, f.z, 0.0f);
        }
    }
    op[k] = pl; od[k] = dep; os[k] = sc; on[k] = ng;
}

// 3x3 median with replicated border (cv::medianBlur, ksize 3, CV_32F)
__global__ void k_median3(const float* __restrict__ src, float* __restrict__ dst, int W, int H) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= W) return;
    float v[9];
    int n = 0;
#pragma unroll
    for (int a = -1; a <= 1; ++a) {
        const int ii = min(max(i + a, 0), H - 1);
#pragma unroll
        for (int b = -1; b <= 1; ++b) {
            const int jj = min(max(j + b, 0), W - 1);
            v[n++] = src[(size_t)ii * W + jj];
        }
    }
    // median of 9 by a fixed exchange network (the value equals any sort's v[4])
#pragma unroll
    for (int p = 0; p <= 4; ++p)
#pragma unroll
        for (int b = p + 1; b < 9; ++b) {
            const float lo = fminf(v[p], v[b]);
            const float hi = fmaxf(v[p], v[b]);
            v[p] = lo;
            v[b] = hi;
        }
    dst[(size_t)i * W + j] = v[4];
}

// PostProcess test (depthmap.cc:1189-1196)
__global__ void k_postfilter(const float* __restrict__ med, float* __restrict__ depth, float* __restrict__ score,
                             int n, float rel_thr, float min_depth, float max_depth, float min_score) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float d = depth[k], m = med[k], sc = score[k];
    if (d == 0.0f || (double)(__builtin_fabsf(d - m) / d) > (double)rel_thr || d < min_depth || d > max_depth ||
        sc <= min_score) {
        depth[k] = 0.0f;
        score[k] = -1.0f;
    } else {
        depth[k] = m;
    }
}

// ---------------------------------------------------------------- host helpers
static int fill_args(pmp_ctx* ctx, const pmp_params* p, KfArgs& a) {
    if (!ctx->views_set) return set_error(ctx, PMP_ESTATE, "pmp_set_views has not been called");
    if (p->patch_half < 3 || p->patch_stride < 1 || p->patch_half > 5 + 5)
        return set_error(ctx, PMP_EINVAL, "patch_half must be >= 3 (propagation reads 3 pixels away)");
    if (p->n_refine < 0 || p->n_refine > kMaxRefine) return set_error(ctx, PMP_EINVAL, "n_refine must be in [0, 8]");
    if (p->agg_mode < 0 || p->agg_mode > 2 || p->edge_mode < 0 || p->edge_mode > 2 || p->coverage_mode < 0 ||
        p->coverage_mode > 1 || p->init_mode < 0 || p->init_mode > 1)
        return set_error(ctx, PMP_EINVAL, "unknown mode value");
    if (!(p->sigma_color > 0.0f) || !(p->sigma_space > 0.0f))
        return set_error(ctx, PMP_EINVAL, "sigma_color and sigma_space must be positive");
    memset(&a, 0, sizeof(a));
    a.W = ctx->W; a.H = ctx->H; a.N = (int)ctx->view_ids.size();
    a.half = p->patch_half; a.stride = p->patch_stride;
    int T = 0;
    for (int k1 = -a.half; k1 <= a.half; k1 += a.stride)
        for (int k2 = -a.half; k2 <= a.half; k2 += a.stride) {
            if (T >= kMaxTaps) return set_error(ctx, PMP_EINVAL, "too many taps (max 121)");
            a.Ws[T++] = (float)exp(-(double)(k1 * k1 + k2 * k2) /
                                   (2.0 * (double)p->sigma_space * (double)p->sigma_space));
        }
    a.T = T;
    if (a.W < 2 * a.half + 8 || a.H < 2 * a.half + 8) return set_error(ctx, PMP_EINVAL, "image too small for the patch");
    a.agg_mode = p->agg_mode; a.edge_mode = p->edge_mode; a.coverage = p->coverage_mode;
    a.init_mode = p->init_mode; a.n_refine = p->n_refine;
    int k = p->topk < 1 ? 1 : p->topk;
    if (k > a.N) k = a.N;
    a.topk = k;
    a.inv_k = (float)(1.0 / (double)k);
    int rc = ensure_tables(ctx, a.W, a.H, p->sigma_color);
    if (rc) return rc;
    a.ref = ctx->images[ctx->ref_id].gray;
    // neighbours packed contiguously (view v at v*W*H) for the current edge mode
    std::vector<uint64_t> key;
    key.push_back(p->edge_mode == PMP_EDGE_WRAP ? 1 : 0);
    for (int v = 0; v < a.N; ++v) {
        key.push_back(ctx->view_ids[v]);
        key.push_back(ctx->images[ctx->view_ids[v]].generation);
    }
    const size_t P = (size_t)a.W * a.H;
    if (key != ctx->packed_key) {
        if (ctx->views_q_cap < (size_t)a.N * P) {
            PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
            if (ctx->views_q) (void)hipFree(ctx->views_q);
            ctx->views_q = nullptr;
            ctx->views_q_cap = 0;
            PMP_HIP(ctx, hipMalloc(&ctx->views_q, (size_t)a.N * P * sizeof(uint32_t)));
            ctx->views_q_cap = (size_t)a.N * P;
        }
        for (int v = 0; v < a.N; ++v) {
            const Image& im = ctx->images[ctx->view_ids[v]];
            dim3 blk(256), grd((im.W + 255) / 256, im.H);
            hipLaunchKernelGGL(k_pack, grd, blk, 0, ctx->stream, im.gray, ctx->views_q + v * P, im.W, im.H,
                               p->edge_mode == PMP_EDGE_WRAP ? 1 : 0);
            PMP_HIP(ctx, hipGetLastError());
        }
        ctx->packed_key = key;
    }
    a.quads = ctx->views_q;
    a.quads_bytes = (uint32_t)((size_t)a.N * P * sizeof(uint32_t));
    a.rowtab = ctx->rowtab; a.coltab = ctx->coltab; a.Wc = ctx->wc;
    derive_poses(ctx->ref_pose, a.N, ctx->view_poses.data(), a.M, a.c);
    a.ku = (float)((double)a.W / (2.0 * M_PI));
    a.hu = (float)(0.5 * a.W);
    a.kv = (float)((double)a.H / M_PI);
    a.hv = (float)(0.5 * a.H);
    a.umax = (p->edge_mode == PMP_EDGE_WRAP) ? (float)a.W : (float)(a.W - 1);
    a.vmax = (float)(a.H - 1);
    a.var_floor = p->var_floor; a.fold_agree = p->fold_agree;
    a.depth_penalty = p->depth_penalty; a.plane_penalty = p->plane_penalty;
    float dr = p->refine_depth_range, nr = p->refine_normal_range;
    for (int t = 0; t < kMaxRefine; ++t) {
        a.dr[t] = dr; a.nr[t] = nr;
        dr = (float)((double)dr * (double)p->refine_depth_decay);
        nr = (float)((double)nr * (double)p->refine_normal_decay);
    }
    a.min_depth = p->min_depth; a.max_depth = p->max_depth;
    a.inv_lo = p->init_inv_depth_min; a.inv_hi = p->init_inv_depth_max;
    a.key0 = (uint32_t)(p->seed & 0xFFFFFFFFu);
    a.key1 = (uint32_t)(p->seed >> 32);
    return PMP_OK;
}

#define PMP_LAUNCH_NV(N, Z, KERNEL, GRID, BLOCK, ...)                                                  \
    switch (N) {                                                                                      \
        case 1: hipLaunchKernelGGL((KERNEL<1, Z>), GRID, BLOCK, 0, ctx->stream, __VA_ARGS__); break;   \
        case 2: hipLaunchKernelGGL((KERNEL<2, Z>), GRID, BLOCK, 0, ctx->stream, __VA_ARGS__); break;   \
        case 3: hipLaunchKernelGGL((KERNEL<3, Z>), GRID, BLOCK, 0, ctx->stream, __VA_ARGS__); break;   \
        case 4: hipLaunchKernelGGL((KERNEL<4, Z>), GRID, BLOCK, 0, ctx->stream, __VA_ARGS__); break;   \
        case 5: hipLaunchKernelGGL((KERNEL<5, Z>), GRID, BLOCK, 0, ctx->stream, __VA_ARGS__); break;   \
        case 6: hipLaunchKernelGGL((KERNEL<6, Z>), GRID, BLOCK, 0, ctx->stream, __VA_ARGS__); break;   \
        case 7: hipLaunchKernelGGL((KERNEL<7, Z>), GRID, BLOCK, 0, ctx->stream, __VA_ARGS__); break;   \
        case 8: hipLaunchKernelGGL((KERNEL<8, Z>), GRID, BLOCK, 0, ctx->stream, __VA_ARGS__); break;   \
        case 9: hipLaunchKernelGGL((KERNEL<9, Z>), GRID, BLOCK, 0, ctx->stream, __VA_ARGS__); break;   \
        case 10: hipLaunchKernelGGL((KERNEL<10, Z>), GRID, BLOCK, 0, ctx->stream, __VA_ARGS__); break; \
        default: return set_error(ctx, PMP_EINVAL, "n_views must be in [1, 10]");                     \
    }

// dispatch on the view count and on zero-outside sampling (PMP_EDGE_ZERO)
#define PMP_DISPATCH_NV(N, KERNEL, GRID, BLOCK, ...)                      \
    if (a.edge_mode == PMP_EDGE_ZERO) {                                  \
        PMP_LAUNCH_NV(N, true, KERNEL, GRID, BLOCK, __VA_ARGS__)         \
    } else {                                                             \
        PMP_LAUNCH_NV(N, false, KERNEL, GRID, BLOCK, __VA_ARGS__)        \
    }

// 1-D grid of 256-thread blocks over tiles of 64 x 4, padded to a multiple of 8
struct TileGrid {
    int nbx, nb;
    dim3 grid() const { return dim3((unsigned)((nb + 7) / 8 * 8)); }
};
static TileGrid tile_grid(int cols, int rows) {
    TileGrid g;
    g.nbx = (cols + 63) / 64;
    g.nb = g.nbx * ((rows + 3) / 4);
    return g;
}

static int launch_init(pmp_ctx* ctx, const KfArgs& a) {
    const TileGrid tg = tile_grid(a.W, a.H);
    const dim3 blk(256), grd = tg.grid();
    const int tp = timing_begin(ctx, kStageInit);
    PMP_DISPATCH_NV(a.N, k_init, grd, blk, a, ctx->work.plane, ctx->work.depth, ctx->work.score, ctx->init.plane,
                    ctx->init.depth, ctx->init.score, tg.nbx, tg.nb);
    timing_end(ctx, tp);
    PMP_HIP(ctx, hipGetLastError());
    return PMP_OK;
}

static int launch_score(pmp_ctx* ctx, const KfArgs& a) {
    const TileGrid tg = tile_grid(a.W, a.H);
    const dim3 blk(256), grd = tg.grid();
    PMP_DISPATCH_NV(a.N, k_score, grd, blk, a, ctx->work.plane, ctx->work.score, tg.nbx, tg.nb);
    PMP_HIP(ctx, hipGetLastError());
    return PMP_OK;
}

static int launch_passes(pmp_ctx* ctx, const KfArgs& a, int it_begin, int it_end) {
    PMP_HIP(ctx, hipMemsetAsync(ctx->counters, 0, 2 * sizeof(unsigned long long), ctx->stream));
    const TileGrid tg = tile_grid((a.W + 1) / 2, a.H);
    const dim3 blk(256), grd = tg.grid();
    for (int it = it_begin; it < it_end; ++it)
        for (int colour = 0; colour < 2; ++colour) {
            const int tp = timing_begin(ctx, kStagePropagate);
            PMP_DISPATCH_NV(a.N, k_propagate, grd, blk, a, ctx->work.plane, ctx->work.depth, ctx->work.score, it,
                            colour, ctx->counters, tg.nbx, tg.nb);
            timing_end(ctx, tp);
            PMP_HIP(ctx, hipGetLastError());
        }
    return PMP_OK;
}

static int launch_finalize(pmp_ctx* ctx, const KfArgs& a, const pmp_params* p, const uint8_t* mask_host) {
    const size_t P = (size_t)a.W * a.H;
    const uint8_t* dmask = nullptr;
    if (mask_host) {
        PMP_HIP(ctx, hipMemcpyAsync(ctx->mask, mask_host, P, hipMemcpyHostToDevice, ctx->stream));
        dmask = ctx->mask;
    } else if (ctx->mask_id >= 0) {
        auto it = ctx->images.find((uint64_t)ctx->mask_id);
        if (it == ctx->images.end()) return set_error(ctx, PMP_ENOTFOUND, "mask image was dropped");
        if (it->second.W != a.W || it->second.H != a.H)
            return set_error(ctx, PMP_EINVAL, "mask image size differs from the reference");
        dmask = it->second.gray;
    }
    dim3 blk(256), grd((a.W + 255) / 256, a.H);
    const int tp = timing_begin(ctx, kStageFinalize);
    hipLaunchKernelGGL(k_copyback, grd, blk, 0, ctx->stream, a, dmask, ctx->init.plane, ctx->init.depth,
                       ctx->init.score, ctx->work.plane, ctx->work.depth, ctx->work.score, ctx->out_plane,
                       ctx->out_depth, ctx->out_score, ctx->out_nghbr, p->min_patch_sd * p->min_patch_sd,
                       (float)(1.0 / (double)a.T));
    PMP_HIP(ctx, hipGetLastError());
    if (p->postprocess) {
        hipLaunchKernelGGL(k_median3, grd, blk, 0, ctx->stream, ctx->out_depth, ctx->tmp0, a.W, a.H);
        hipLaunchKernelGGL(k_median3, grd, blk, 0, ctx->stream, ctx->tmp0, ctx->tmp1, a.W, a.H);
        hipLaunchKernelGGL(k_postfilter, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, ctx->stream, ctx->tmp1,
                           ctx->out_depth, ctx->out_score, (int)P, p->median_rel_thr, p->min_depth, p->max_depth,
                           p->min_score);
        PMP_HIP(ctx, hipGetLastError());
    }
    timing_end(ctx, tp);
    return PMP_OK;
}

}  // namespace pmp

using namespace pmp;

extern "C" {

int pmp_set_views(pmp_ctx* ctx, uint64_t ref_id, const pmp_pose* ref_pose, int32_t n_views, const uint64_t* view_ids,
                  const pmp_pose* view_poses) {
    if (!ctx || !ref_pose || (n_views > 0 && (!view_ids || !view_poses))) return PMP_EINVAL;
    if (n_views < 1 || n_views > kMaxViews) return set_error(ctx, PMP_EINVAL, "n_views must be in [1, 10]");
    auto it = ctx->images.find(ref_id);
    if (it == ctx->images.end()) return set_error(ctx, PMP_ENOTFOUND, "unknown reference image id");
    const int W = it->second.W, H = it->second.H;
    for (int v = 0; v < n_views; ++v) {
        auto iv = ctx->images.find(view_ids[v]);
        if (iv == ctx->images.end()) return set_error(ctx, PMP_ENOTFOUND, "unknown neighbour image id");
        if (iv->second.W != W || iv->second.H != H)
            return set_error(ctx, PMP_EINVAL, "neighbour images must have the reference's size");
    }
    int rc = ensure_workspace(ctx, W, H);
    if (rc) return rc;
    ctx->ref_id = ref_id;
    ctx->view_ids.assign(view_ids, view_ids + n_views);
    memcpy(ctx->ref_pose, ref_pose->R, 9 * sizeof(double));
    memcpy(ctx->ref_pose + 9, ref_pose->t, 3 * sizeof(double));
    ctx->view_poses.resize(12 * (size_t)n_views);
    for (int v = 0; v < n_views; ++v) {
        memcpy(&ctx->view_poses[12 * v], view_poses[v].R, 9 * sizeof(double));
        memcpy(&ctx->view_poses[12 * v + 9], view_poses[v].t, 3 * sizeof(double));
    }
    ctx->W = W; ctx->H = H;
    ctx->views_set = true;
    return PMP_OK;
}

int pmp_set_mask(pmp_ctx* ctx, int64_t mask_id) {
    if (!ctx) return PMP_EINVAL;
    if (mask_id >= 0 && ctx->images.find((uint64_t)mask_id) == ctx->images.end())
        return set_error(ctx, PMP_ENOTFOUND, "unknown mask image id");
    ctx->mask_id = mask_id < 0 ? -1 : mask_id;
    return PMP_OK;
}

int pmp_random_initialization(pmp_ctx* ctx, const pmp_params* params) {
    if (!ctx || !params) return PMP_EINVAL;
    KfArgs a;
    int rc = fill_args(ctx, params, a);
    if (rc) return rc;
    return launch_init(ctx, a);
}

int pmp_score_plane_depth(pmp_ctx* ctx, const pmp_params* params) {
    if (!ctx || !params) return PMP_EINVAL;
    KfArgs a;
    int rc = fill_args(ctx, params, a);
    if (rc) return rc;
    return launch_score(ctx, a);
}

int pmp_red_black_pass(pmp_ctx* ctx, const pmp_params* params, int32_t it_begin, int32_t it_end) {
    if (!ctx || !params || it_begin < 0 || it_end < it_begin) return PMP_EINVAL;
    KfArgs a;
    int rc = fill_args(ctx, params, a);
    if (rc) return rc;
    return launch_passes(ctx, a, it_begin, it_end);
}

int pmp_finalize(pmp_ctx* ctx, const pmp_params* params, const uint8_t* mask) {
    if (!ctx || !params) return PMP_EINVAL;
    KfArgs a;
    int rc = fill_args(ctx, params, a);
    if (rc) return rc;
    return launch_finalize(ctx, a, params, mask);
}

int pmp_densify(pmp_ctx* ctx, const pmp_params* params, const uint8_t* mask) {
    if (!ctx || !params || params->iterations < 0) return PMP_EINVAL;
    KfArgs a;
    int rc = fill_args(ctx, params, a);
    if (rc) return rc;
    if ((rc = launch_init(ctx, a))) return rc;
    if ((rc = launch_passes(ctx, a, 0, params->iterations))) return rc;
    if ((rc = launch_finalize(ctx, a, params, mask))) return rc;
    ctx->densify_calls++;
    return PMP_OK;
}

int pmp_set_maps(pmp_ctx* ctx, const float* plane4, const float* depth, const float* score) {
    if (!ctx) return PMP_EINVAL;
    if (!ctx->views_set) return set_error(ctx, PMP_ESTATE, "pmp_set_views has not been called");
    const size_t P = (size_t)ctx->W * ctx->H;
    if (plane4) PMP_HIP(ctx, hipMemcpyAsync(ctx->work.plane, plane4, P * 16, hipMemcpyHostToDevice, ctx->stream));
    if (depth) PMP_HIP(ctx, hipMemcpyAsync(ctx->work.depth, depth, P * 4, hipMemcpyHostToDevice, ctx->stream));
    if (score) PMP_HIP(ctx, hipMemcpyAsync(ctx->work.score, score, P * 4, hipMemcpyHostToDevice, ctx->stream));
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PMP_OK;
}

int pmp_get_maps(pmp_ctx* ctx, int32_t which, float* plane4, float* depth, float* score, int32_t* nghbr) {
    if (!ctx) return PMP_EINVAL;
    if (!ctx->views_set) return set_error(ctx, PMP_ESTATE, "pmp_set_views has not been called");
    const size_t P = (size_t)ctx->W * ctx->H;
    const void *sp, *sd, *ss;
    if (which == PMP_MAPS_WORK) { sp = ctx->work.plane; sd = ctx->work.depth; ss = ctx->work.score; }
    else if (which == PMP_MAPS_INIT) { sp = ctx->init.plane; sd = ctx->init.depth; ss = ctx->init.score; }
    else if (which == PMP_MAPS_FINAL) { sp = ctx->out_plane; sd = ctx->out_depth; ss = ctx->out_score; }
    else return set_error(ctx, PMP_EINVAL, "unknown map set");
    if (plane4) PMP_HIP(ctx, hipMemcpyAsync(plane4, sp, P * 16, hipMemcpyDeviceToHost, ctx->stream));
    if (depth) PMP_HIP(ctx, hipMemcpyAsync(depth, sd, P * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (score) PMP_HIP(ctx, hipMemcpyAsync(score, ss, P * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (nghbr) {
        if (which != PMP_MAPS_FINAL) return set_error(ctx, PMP_EINVAL, "nghbr exists only for the final maps");
        PMP_HIP(ctx, hipMemcpyAsync(nghbr, ctx->out_nghbr, P * 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PMP_OK;
}

int pmp_eval_counts(pmp_ctx* ctx, int64_t* evals, int64_t* candidates) {
    if (!ctx) return PMP_EINVAL;
    unsigned long long h[2] = {0, 0};
    if (ctx->counters) {
        PMP_HIP(ctx, hipMemcpyAsync(h, ctx->counters, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        PMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (evals) *evals = (int64_t)h[0];
    if (candidates) *candidates = (int64_t)h[1];
    return PMP_OK;
}

}  // extern "C"
