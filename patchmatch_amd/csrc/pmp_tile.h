// pmp_tile.h -- per-tile and per-pixel device code shared by the one-keyframe
// kernels (pmp_estimator.hip) and the work-queue kernels over many keyframes
// (pmp_batch.hip).
#pragma once
#include "pmp_internal.h"
#include "pmp_math.h"
#include "pmp_score.h"

namespace pmp {

// Colour-weight table into LDS; every thread of the (1-D) block takes part.
__device__ __forceinline__ void load_wc(const float* wc, float* sWc) {
    for (int k = threadIdx.x; k < 256; k += blockDim.x) sWc[k] = wc[k];
    __syncthreads();
}

__device__ __forceinline__ void count_add(unsigned long long* ctr, unsigned int v) {
    // one atomic per wave (all 64 lanes active; per-lane counts are < 2^16)
    unsigned int s = v;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(ctr, (unsigned long long)s);
}

// Workgroups are dispatched to the 8 XCDs round-robin by linear block id.  The
// grid is padded to a multiple of 8 and block b is given tile
// (b mod 8) * (nb8/8) + b/8, so each XCD walks one contiguous band of tiles
// and its L2 holds that band's neighbour rows.  Returns -1 for padding blocks.
__device__ __forceinline__ int xcd_tile(int b, int nb8, int nb) {
    const int t = (b & 7) * (nb8 >> 3) + (b >> 3);
    return t < nb ? t : -1;
}

// Random hypothesis, initial score and snapshot of pixel (i, j)
// (RandomInitialization, PatchMatchCuda.cu:210-270, depthmap.cc:656-663).
// sv/ld: per-view score scratch of the calling thread (NV == 0 only); tp:
// where the reference taps come from (GlobalTaps or a staged LdsTaps).
template <int NV, bool kZero, class Taps>
__device__ __forceinline__ void init_pixel(const KfArgs& a, const Taps& tp, const float* sWc, int i, int j,
                                           float4* __restrict__ plane, float* __restrict__ depth,
                                           float* __restrict__ score, float4* __restrict__ iplane,
                                           float* __restrict__ idepth, float* __restrict__ iscore,
                                           float* sv = nullptr, int ld = 0) {
    float p0 = 0.0f, p1 = 0.0f, p2 = 0.0f, dep = 0.0f;
    if (in_init_grid(a, i, j)) random_plane(a, i, j, p0, p1, p2, dep);
    float sc = __uint_as_float(0xFFFFFFFFu);   // cudaMemset(score, 0xFF) (depthmap.cc:657)
    if (in_score_grid(a, i, j)) {
        RefStat rs = ref_stats(a, sWc, i, j);
        sc = score_plane_t<NV, kZero>(a, tp, sWc, rs, i, j, p0, p1, p2, sv, ld);
    }
    const size_t k = (size_t)i * a.W + j;
    const float4 pl = make_float4(p0, p1, p2, 0.0f);
    plane[k] = pl; depth[k] = dep; score[k] = sc;
    iplane[k] = pl; idepth[k] = dep; iscore[k] = sc;
}

// ---- red-black pass
// A tile is 2R rows x 32 same-colour pixels (64 columns); thread t of the
// 64R-thread block updates row i = 2R by + t / 32, column
// j = 2 (32 bx + t % 32) + ((i + colour) & 1), i.e. every pixel of colour
// (i+j)&1 == colour once (each wave covers two rows).  Candidates are at odd
// distance (1 or 3), so they have the other colour and are not written by
// this pass.
//
// Propagation candidates are scored by whichever threads of the block are
// free, in packed sweeps, and their raw scores handed back to the owning
// threads, which apply the penalty and the acceptance test in candidate order:
//   1. every pixel tests its 8 candidates against its state at entry (the
//      exact pre-rejection below); the (candidate, pixel) pairs that pass are
//      listed candidate-major and scored in one sweep over the block's threads;
//   2. each pixel then walks its candidates in order against its running state
//      (the test is repeated: an acceptance changes the plane and the score it
//      is tested against).  A candidate that passes there but did not pass at
//      entry has no score yet: the pixel stops and lists every later candidate
//      that can still pass at all (1 - f_d above its current score: the plane
//      factor is at most 1 and the score only grows), and a second sweep scores
//      those of all such pixels before the walks go on.  After it no walk
//      stops again, so a tile takes at most two sweeps.
// A candidate's raw score depends only on its plane and the pixel, so where
// and when it is computed does not matter: each pixel's sequence of decisions
// is that of the one-thread-per-pixel loop, and the results are identical.
// Scores computed for candidates that an acceptance makes unnecessary are
// spent work (counted as evaluations).
// NVS: per-view score scratch rows (kMaxViews for the run-time view count, 1
// otherwise).
// srow/scol/sref: the tile's rows and columns of the bearing tables and of
// the reference image, with the patch halo (LdsTaps).
// res[c * 64R + t]: raw score of candidate c of thread t's pixel; job: the
// sweep's (candidate << 9 | thread) list.
template <int R, int NVS = 1>
struct PropShared {
    float sWc[256];
    float sv[NVS][64 * R];
    float res[8 * 64 * R];
    float4 rs[64 * R];
    float ic[64 * R];
    uint16_t job[8 * 64 * R];
    float2 srow[64 * R / kTileSame + 2 * kMaxHalf];
    float2 scol[kTileCols];
    uint8_t sref[(64 * R / kTileSame + 2 * kMaxHalf) * kTileCols];
    int cnt[8 * R];
    int item;
};

// Offset (di, dj) of propagation candidate c: (0,+1) (+1,0) (0,-1) (-1,0)
// (0,+3) (+3,0) (0,-3) (-3,0) (PatchMatchCuda.cu:305)
__device__ __forceinline__ void cand_offset(int c, int& di, int& dj) {
    const int dist = (c < 4) ? 1 : 3, dir = c & 3;
    di = (dir == 1) ? dist : (dir == 3) ? -dist : 0;
    dj = (dir == 0) ? dist : (dir == 2) ? -dist : 0;
}

// 1 - f_d of candidate plane pn: the largest penalty factor it can have for a
// pixel with bearing b and entry depth d0 (PatchMatchCuda.cu:323-326), whatever
// the running plane.
__device__ __forceinline__ float cand_bound(const KfArgs& a, float4 pn, float bx, float by, float bz, float d0,
                                           float inv_d0) {
    const float s = fma_(pn.z, bz, fma_(pn.y, by, pn.x * bx));
    const float dn = rcp_pos(clampf(abs_(s), 1e-6f, 1e30f));
    const float rel = (d0 - dn) * inv_d0;
    const float fd = exp_((rel * rel) * -a.depth_penalty);
    return 1.0f - fd;
}

// Penalty and acceptance precondition of candidate plane pn for a pixel with
// bearing b, entry depth d0 and running plane h and score sc
// (PatchMatchCuda.cu:323-332).  acosf(prod/norm) is NaN, i.e. the candidate
// is rejected, unless |prod| <= norm; |score| <= 1, so the penalised score
// cannot exceed f, and f <= sc rejects it before it is scored.
__device__ __forceinline__ bool cand_test(const KfArgs& a, float4 pn, float bx, float by, float bz, float d0,
                                          float inv_d0, float h0, float h1, float h2, float sc, float& f) {
    const float gd = cand_bound(a, pn, bx, by, bz, d0, inv_d0);
    const float prod = fma_(h2, pn.z, fma_(h1, pn.y, h0 * pn.x));
    const float ex = h0 - pn.x, ey = h1 - pn.y, ez = h2 - pn.z;
    const float nrm = sqrt_(fma_(ez, ez, fma_(ey, ey, ex * ex)));
    f = 0.0f;
    if (nrm > 0.0f && abs_(prod) <= nrm) {
        const float ratio = clampf(prod * rcp_pos(clampf(nrm, 1e-30f, 1e30f)), -1.0f, 1.0f);
        const float ac = acos_(ratio);
        const float fp = exp_((ac * ac) * -a.plane_penalty);
        f = gd * (1.0f - fp);
        return f > sc;
    }
    return false;
}

// A propagation tile is TR = 64 R / kTileSame rows by 2 kTileSame columns
// (kTileSame same-colour pixels per row; each wave covers 64 / kTileSame
// rows).  Taller, narrower tiles than one row per wave shrink the views' cache
// footprint of a tile's patches.  stage_tile copies the rows
// [tby TR - half, tby TR + TR + half) and columns [2 kTileSame tbx - half,
// 2 kTileSame (tbx + 1) + half) of the reference image and of the bearing
// tables into sh (coordinates outside the image are clamped; only interior
// pixels, whose taps lie inside the image, read them).  Waits for the previous
// tile's readers first.
template <int R, int NVS>
__device__ __forceinline__ LdsTaps stage_tile(const KfArgs& a, PropShared<R, NVS>& sh, int tbx, int tby) {
    constexpr int NT = 64 * R;
    constexpr int TR = NT / kTileSame;   // tile rows
    const int r0 = tby * TR - a.half, c0 = 2 * kTileSame * tbx - a.half;
    const int nr = TR + 2 * a.half, nc = 2 * kTileSame + 2 * a.half;
    __syncthreads();
    for (int e = threadIdx.x; e < nc; e += NT) {
        const int x = min(max(c0 + e, 0), a.W - 1);
        sh.scol[e] = a.coltab[x];
    }
    for (int e = threadIdx.x; e < nr; e += NT) {
        const int y = min(max(r0 + e, 0), a.H - 1);
        sh.srow[e] = a.rowtab[y];
    }
    for (int r = 0; r < nr; ++r) {
        const int y = min(max(r0 + r, 0), a.H - 1);
        for (int e = threadIdx.x; e < nc; e += NT) {
            const int x = min(max(c0 + e, 0), a.W - 1);
            sh.sref[r * kTileCols + e] = a.ref[(size_t)y * a.W + x];
        }
    }
    __syncthreads();
    return LdsTaps{sh.srow, sh.scol, sh.sref, r0, c0};
}

// One tile of one red-black pass.  sh.sWc holds a.Wc; rot (a multiple of 64)
// is the slot offset where the packed jobs start.
template <int NV, bool kZero, int R, int NVS>
__device__ __forceinline__ void propagate_tile(const KfArgs& a, PropShared<R, NVS>& sh, float4* __restrict__ plane,
                                               float* __restrict__ depth, float* __restrict__ score, int it,
                                               int colour, int tbx, int tby, int rot, unsigned int& n_eval,
                                               unsigned int& n_cand) {
    constexpr int NT = 64 * R;
    const float* sWc = sh.sWc;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int TR = NT / kTileSame;   // tile rows
    const int i = tby * TR + tid / kTileSame;
    const int j = 2 * (tbx * kTileSame + tid % kTileSame) + ((i + colour) & 1);
    const bool active = i < a.H && j < a.W && in_prop_grid(a, i, j);
    const int W = a.W;
    const size_t k = active ? (size_t)i * W + j : 0;

    const LdsTaps tp = stage_tile(a, sh, tbx, tby);
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
    sh.rs[tid] = make_float4(rs.mx, rs.inv, rs.mxc, rs.vx);
    sh.ic[tid] = rs.Ic;

    // 1. propagation (PatchMatchCuda.cu:311-338), in the order of cand_offset
    unsigned int want = 0;     // bit c: candidate c is to be scored in the next sweep
    if (active) {
        n_cand += 8;
#pragma unroll 1
        for (int c = 0; c < 8; ++c) {
            int di, dj;
            cand_offset(c, di, dj);
            const float4 pn = plane[(size_t)(i + di) * W + (j + dj)];
            float f;
            if (cand_test(a, pn, bx, by, bz, d0, inv_d0, h0, h1, h2, sc, f)) want |= 1u << c;
        }
    }
    unsigned int listed = want;   // bit c: candidate c has a score (or will after the sweep)
    int cpos = 0;                 // the next candidate of this pixel's walk
#pragma unroll 1
    for (;;) {
        // the sweep's list, candidate-major: per (candidate, wave) counts, then offsets
#pragma unroll 1
        for (int c = 0; c < 8; ++c) {
            const unsigned long long bal = __ballot((want >> c) & 1u);
            if (lane == 0) sh.cnt[c * R + wave] = __popcll(bal);
        }
        __syncthreads();
        int J = 0;
#pragma unroll 1
        for (int c = 0; c < 8; ++c) {
            int woff = 0, tot = 0;
#pragma unroll
            for (int w = 0; w < R; ++w) {
                const int cw = sh.cnt[c * R + w];
                woff += (w < wave) ? cw : 0;
                tot += cw;
            }
            const bool on = (want >> c) & 1u;
            const unsigned long long bal = __ballot(on);
            if (on) sh.job[J + woff + __popcll(bal & ((1ull << lane) - 1ull))] = (uint16_t)((c << 9) | tid);
            J += tot;
        }
        if (J == 0) break;   // the same in every thread
        __syncthreads();
        // score the list: job g by thread (g + rot) mod NT
#pragma unroll 1
        for (int g = (tid - rot) & (NT - 1); g < J; g += NT) {
            const unsigned int e = sh.job[g];
            const int c = (int)(e >> 9), jl = (int)(e & 511u);
            const int il = tby * TR + jl / kTileSame;
            const int jc = 2 * (tbx * kTileSame + jl % kTileSame) + ((il + colour) & 1);
            int di, dj;
            cand_offset(c, di, dj);
            const float4 pn = plane[(size_t)(il + di) * W + (jc + dj)];
            const float4 q = sh.rs[jl];
            RefStat r2;
            r2.Ic = sh.ic[jl]; r2.mx = q.x; r2.inv = q.y; r2.mxc = q.z; r2.vx = q.w;
            sh.res[c * NT + jl] = score_plane_t<NV, kZero>(a, tp, sWc, r2, il, jc, pn.x, pn.y, pn.z, &sh.sv[0][tid], NT);
            ++n_eval;
        }
        __syncthreads();
        // walk the candidates in order against the running state, up to the
        // first one that passes but has no score
        want = 0;
        if (active) {
            while (cpos < 8) {
                int di, dj;
                cand_offset(cpos, di, dj);
                const float4 pn = plane[(size_t)(i + di) * W + (j + dj)];
                float f;
                if (cand_test(a, pn, bx, by, bz, d0, inv_d0, h0, h1, h2, sc, f)) {
                    if (!((listed >> cpos) & 1u)) break;
                    const float sh_ = sh.res[cpos * NT + tid] * f;
                    if (sh_ > sc) { h0 = pn.x; h1 = pn.y; h2 = pn.z; sc = sh_; }
                }
                ++cpos;
            }
            if (cpos < 8) {
                // stopped: list this and every later unscored candidate that can still pass
#pragma unroll 1
                for (int c = cpos; c < 8; ++c) {
                    if ((listed >> c) & 1u) continue;
                    int di, dj;
                    cand_offset(c, di, dj);
                    const float4 pn = plane[(size_t)(i + di) * W + (j + dj)];
                    if (c == cpos || cand_bound(a, pn, bx, by, bz, d0, inv_d0) > sc) want |= 1u << c;
                }
                listed |= want;
            }
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
            const float sh_ = score_plane_t<NV, kZero>(a, tp, sWc, rs, i, j, q0, q1, q2, &sh.sv[0][tid], NT);
            ++n_eval;
            ++n_cand;
            if (sh_ > sc) { c0 = q0; c1 = q1; c2 = q2; sc = sh_; dcur = dh; }
        }
        score[k] = sc;
        plane[k] = make_float4(c0, c1, c2, 0.0f);
        depth[k] = depth_of_plane(bx, by, bz, c0, c1, c2);
    }
}

// ---- finalize
// Copy-back of ComputePatchMatch: masked pixels keep the AssignMatrices values
// (depth 0, plane 0, score -1, nghbr 0); unmasked pixels take the maps after
// RandomInitialization (depthmap.cc:673-687) unless they are inside the window
// [ps-2, dim-ps-2) (depthmap.cc:731-732), where they take the final maps, or
// 0 when the patch variance is below min_patch_sd^2 (depthmap.cc:739-743).
__device__ __forceinline__ void copyback_pixel(const KfArgs& a, const uint8_t* __restrict__ mask,
                                               const float4* __restrict__ ip, const float* __restrict__ id,
                                               const float* __restrict__ is, const float4* __restrict__ fp,
                                               const float* __restrict__ fd, const float* __restrict__ fs,
                                               float4* __restrict__ op, float* __restrict__ od,
                                               float* __restrict__ os, int32_t* __restrict__ on, float min_var,
                                               float inv_T, int i, int j) {
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
__device__ __forceinline__ void median3_pixel(const float* __restrict__ src, float* __restrict__ dst, int W, int H,
                                              int i, int j) {
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

// PostProcess test (depthmap.cc:1189-1196): the float ratio |d - m| / d against
// the double threshold (fabs of a float is float there, 0.05 a double literal)
__device__ __forceinline__ void postfilter_pixel(const float* __restrict__ med, float* __restrict__ depth,
                                                 float* __restrict__ score, size_t k, double rel_thr,
                                                 float min_depth, float max_depth, float min_score) {
    const float d = depth[k], m = med[k], sc = score[k];
    if (d == 0.0f || (double)(__builtin_fabsf(d - m) / d) > rel_thr || d < min_depth || d > max_depth ||
        sc <= min_score) {
        depth[k] = 0.0f;
        score[k] = -1.0f;
    } else {
        depth[k] = m;
    }
}

}  // namespace pmp
