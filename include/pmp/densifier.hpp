// densifier.hpp -- C++ facade over the pmp C-ABI with the method set of the
// reference's dense:: classes, for a drop-in into the OpenVSLAM fork's
// densification thread (depthmap_computation_module.cc:205-472) and point-cloud
// thread (dense_pointcloud_module.cc:56-182).
//
//   pmp::DepthmapEstimator  <- dense::DepthmapEstimator (depthmap.h:122-202)
//   pmp::DepthmapCleaner    <- dense::DepthmapCleaner   (depthmap.h:212-241)
//   pmp::DepthmapPruner     <- dense::DepthmapPruner    (depthmap.h:254-291)
//
// Differences of form, not of behaviour: images, masks and maps are plain
// row-major buffers instead of cv::Mat, rotations are row-major double[9];
// errors throw pmp::Error (the reference printed CUDA errors and continued).
// Header-only; link with libpmp.so.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "pmp/pmp.h"

namespace pmp {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(pmp_ctx* ctx, int rc, const char* what) {
    if (rc != PMP_OK) throw Error(rc, std::string(what) + ": " + (ctx ? pmp_last_error(ctx) : "no context"));
}

// One shared device context per facade object set (not thread-safe, like the
// reference classes).
class Context {
public:
    explicit Context(int device = 0) {
        int rc = pmp_create(&ctx_, device);
        if (rc != PMP_OK) throw Error(rc, "pmp_create failed (is a GPU visible?)");
    }
    ~Context() { pmp_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    pmp_ctx* get() const { return ctx_; }

private:
    pmp_ctx* ctx_ = nullptr;
};

// DepthmapEstimatorResult (depthmap.h:71-79)
struct DepthmapEstimatorResult {
    int rows = 0, cols = 0;
    std::vector<float> depth;    // CV_32F
    std::vector<double> plane;   // CV_64FC3 (3 per pixel)
    std::vector<float> score;    // CV_32F
    std::vector<int32_t> nghbr;  // CV_32S
    double R[9] = {0};
    double t[3] = {0};
};

// DepthmapCleanerResult (depthmap.h:204-210)
struct DepthmapCleanerResult {
    int rows = 0, cols = 0;
    std::vector<float> cleaned_depth;
    double R[9] = {0};
    double t[3] = {0};
};

// DepthmapPrunerResult (depthmap.h:243-252)
struct DepthmapPrunerResult {
    int rows = 0, cols = 0;
    std::vector<float> pruned_depth;
    std::vector<double> points3d;    // 3 per point, world frame
    std::vector<uint8_t> colors;     // 3 per point, BGR of the thumbnail
    std::vector<int32_t> keypts;     // (row, col) per point
    double R[9] = {0};
    double t[3] = {0};
};

inline pmp_pose make_pose(const double* R, const double* t) {
    pmp_pose p;
    std::memcpy(p.R, R, sizeof(p.R));
    std::memcpy(p.t, t, sizeof(p.t));
    return p;
}

class DepthmapEstimator {
public:
    explicit DepthmapEstimator(Context& ctx) : ctx_(ctx) {
        pmp_default_params(&params_);
        params_.patch_half = patch_size_ - 2;
    }

    // AddView (depthmap.cc:218-236): rotation/translation world->camera,
    // grey image and mask (W*H bytes, row-major), the keyframe's landmarks
    // (kept for warm-start seeding), depth range.
    void AddView(const double R[9], const double t[3], const uint8_t* img, const uint8_t* mask, int width,
                 int height, const std::vector<double>& landmarks_xyz, double maxd, double mind) {
        View v;
        v.pose = make_pose(R, t);
        v.width = width;
        v.height = height;
        v.id = next_id_++;
        v.mask.assign(mask, mask + (size_t)width * height);
        v.landmarks = landmarks_xyz;
        v.maxd = maxd;
        v.mind = mind;
        check(ctx_.get(), pmp_image_put(ctx_.get(), v.id, img, width, height, width, 0), "pmp_image_put");
        views_.push_back(std::move(v));
    }

    bool IsReadyForCompute() const { return (int)views_.size() >= min_images_compute_; }
    bool SetMinImagesCompute(int n) {
        bool ok = n >= 2;
        min_images_compute_ = ok ? n : 2;
        return ok;
    }
    void SetDepthRange(double min_depth, double max_depth, int /*num_depth_planes*/) {
        params_.min_depth = (float)min_depth;
        params_.max_depth = (float)max_depth;
    }
    void SetPatchMatchIterations(int n) { params_.iterations = n; }
    // patchhalf = patch_size_ - 2 (depthmap.cc:666,725)
    void SetPatchSize(int size) {
        patch_size_ = size;
        params_.patch_half = size - 2;
    }
    void SetMinPatchSD(float sd) { params_.min_patch_sd = sd; }
    void SetMinScore(double s) { params_.min_score = (float)s; }
    pmp_params& params() { return params_; }

    // ComputePatchMatch (depthmap.cc:541-806): the centre of the window is the
    // reference (swapCenter, :835-845), the others its neighbours in the
    // swapped order; depth range from the centre view.
    void ComputePatchMatch(DepthmapEstimatorResult* result) {
        if (!IsReadyForCompute()) throw Error(PMP_ESTATE, "not enough views");
        std::vector<size_t> order(views_.size());
        for (size_t k = 0; k < order.size(); ++k) order[k] = k;
        std::swap(order[0], order[views_.size() / 2]);
        const View& ref = views_[order[0]];
        std::vector<uint64_t> ids;
        std::vector<pmp_pose> poses;
        for (size_t k = 1; k < order.size(); ++k) {
            ids.push_back(views_[order[k]].id);
            poses.push_back(views_[order[k]].pose);
        }
        pmp_params p = params_;
        p.min_depth = (float)ref.mind;
        p.max_depth = (float)ref.maxd;
        pmp_ctx* c = ctx_.get();
        check(c, pmp_set_views(c, ref.id, &ref.pose, (int32_t)ids.size(), ids.data(), poses.data()), "pmp_set_views");
        // RandomInitialization, then the pre-initialisation (depthmap.cc:656-691):
        // the reference's landmarks, then the memorised previous result
        check(c, pmp_random_initialization(c, &p), "pmp_random_initialization");
        if (!ref.landmarks.empty())
            check(c, pmp_preinit_landmarks(c, &p, (int64_t)(ref.landmarks.size() / 3), ref.landmarks.data(),
                                           ref.mask.data(), nullptr),
                  "pmp_preinit_landmarks");
        check(c, pmp_preinit_warp(c, &p, ref.mask.data(), nullptr), "pmp_preinit_warp");
        check(c, pmp_red_black_pass(c, &p, 0, p.iterations), "pmp_red_black_pass");
        check(c, pmp_finalize(c, &p, ref.mask.data()), "pmp_finalize");
        const size_t P = (size_t)ref.width * ref.height;
        std::vector<float> plane4(4 * P);
        result->rows = ref.height;
        result->cols = ref.width;
        result->depth.resize(P);
        result->score.resize(P);
        result->nghbr.resize(P);
        check(ctx_.get(),
              pmp_get_maps(ctx_.get(), PMP_MAPS_FINAL, plane4.data(), result->depth.data(), result->score.data(),
                           result->nghbr.data()),
              "pmp_get_maps");
        result->plane.resize(3 * P);
        for (size_t k = 0; k < P; ++k)
            for (int c = 0; c < 3; ++c) result->plane[3 * k + c] = (double)plane4[4 * k + c];
        std::memcpy(result->R, ref.pose.R, sizeof(result->R));
        std::memcpy(result->t, ref.pose.t, sizeof(result->t));
        check(c, pmp_memorize_result(c), "pmp_memorize_result");   // MemorizeResult (:768)
    }

    // PopHeadUnlessOne (depthmap.cc:241-254)
    int PopHeadUnlessOne() {
        if (views_.size() > 1) {
            pmp_image_drop(ctx_.get(), views_.front().id);
            views_.erase(views_.begin());
        }
        return (int)views_.size();
    }

    // reset (depthmap.cc:256-280): the memorised result is cleared too
    void reset() {
        for (auto& v : views_) pmp_image_drop(ctx_.get(), v.id);
        views_.clear();
        pmp_forget_result(ctx_.get());
    }

private:
    struct View {
        pmp_pose pose;
        int width = 0, height = 0;
        uint64_t id = 0;
        std::vector<uint8_t> mask;
        std::vector<double> landmarks;
        double maxd = 0, mind = 0;
    };
    Context& ctx_;
    pmp_params params_;
    std::vector<View> views_;
    int patch_size_ = 7;
    int min_images_compute_ = 2;
    uint64_t next_id_ = 1ull << 40;
};

class DepthmapCleaner {
public:
    explicit DepthmapCleaner(Context& ctx) : ctx_(ctx) {}
    void SetSameDepthThreshold(float t) { thr_ = t; }
    void SetMinConsistentViews(int n) { min_views_ = n; }
    void SetDepthQueueSize(int n) { queue_ = n; }
    // Optional tests of pmp_clean_ex (no counterpart in the reference; <= 0: off):
    // round trip into the centre map within px pixels, normals within deg degrees.
    void SetMaxReprojectionPx(float px) { max_px_ = px; }
    void SetMaxNormalAngleDeg(float deg) { max_deg_ = deg; }
    bool IsReadyForCompute() const { return (int)results_.size() >= queue_; }
    void AddView(const DepthmapEstimatorResult& r) { results_.push_back(r); }
    // Clean (depthmap.cc:1320-1377): the centre result is cleaned against the others
    void Clean(DepthmapCleanerResult* out) {
        if (!IsReadyForCompute() || results_.empty()) return;
        std::vector<size_t> order(results_.size());
        for (size_t k = 0; k < order.size(); ++k) order[k] = k;
        std::swap(order[0], order[results_.size() / 2]);
        std::vector<const float*> maps;
        std::vector<pmp_pose> poses;
        for (size_t k : order) {
            maps.push_back(results_[k].depth.data());
            poses.push_back(make_pose(results_[k].R, results_[k].t));
        }
        const DepthmapEstimatorResult& c = results_[order[0]];
        out->rows = c.rows;
        out->cols = c.cols;
        out->cleaned_depth.assign((size_t)c.rows * c.cols, 0.0f);
        if (max_px_ > 0.0f || max_deg_ > 0.0f) {
            // planes back to the 4-float layout of the maps (they were floats before the widening)
            std::vector<std::vector<float>> p4(order.size());
            std::vector<const float*> planes;
            for (size_t q = 0; q < order.size(); ++q) {
                const std::vector<double>& pl = results_[order[q]].plane;
                p4[q].assign(pl.size() / 3 * 4, 0.0f);
                for (size_t k = 0; k < pl.size() / 3; ++k)
                    for (int e = 0; e < 3; ++e) p4[q][4 * k + e] = (float)pl[3 * k + e];
                planes.push_back(p4[q].data());
            }
            check(ctx_.get(),
                  pmp_clean_ex(ctx_.get(), c.cols, c.rows, (int32_t)maps.size(), maps.data(), planes.data(),
                               poses.data(), thr_, min_views_, max_px_, max_deg_, out->cleaned_depth.data()),
                  "pmp_clean_ex");
        } else {
            check(ctx_.get(),
                  pmp_clean(ctx_.get(), c.cols, c.rows, (int32_t)maps.size(), maps.data(), poses.data(), thr_,
                            min_views_, out->cleaned_depth.data()),
                  "pmp_clean");
        }
        std::memcpy(out->R, c.R, sizeof(out->R));
        std::memcpy(out->t, c.t, sizeof(out->t));
    }
    int PopHeadUnlessOne() {
        if (results_.size() > 1) results_.erase(results_.begin());
        return (int)results_.size();
    }
    void reset() { results_.clear(); }

private:
    Context& ctx_;
    std::vector<DepthmapEstimatorResult> results_;
    float thr_ = 0.01f;    // depthmap.cc:1211
    int min_views_ = 2;
    int queue_ = 1;
    float max_px_ = 0.0f, max_deg_ = 0.0f;
};

class DepthmapPruner {
public:
    explicit DepthmapPruner(Context& ctx) : ctx_(ctx) {}
    void SetSameDepthThreshold(float t) { thr_ = t; }
    void SetMinViews(int n) { min_views_ = n; }
    void SetDepthQueueSize(int n) { queue_ = n; }
    bool IsReadyForCompute() const { return pmp_pruner_size(ctx_.get()) >= queue_; }
    // AddView (depthmap.cc:1405-1423); bgr: W*H*3 thumbnail
    void AddView(const DepthmapCleanerResult& r, const uint8_t* bgr) {
        pmp_pose p = make_pose(r.R, r.t);
        check(ctx_.get(), pmp_pruner_add_view(ctx_.get(), r.cleaned_depth.data(), &p, bgr, r.cols, r.rows),
              "pmp_pruner_add_view");
        poses_.push_back(p);
        dims_.push_back({r.rows, r.cols});
    }
    // Prune (depthmap.cc:1535-1696)
    void Prune(DepthmapPrunerResult* out) {
        if (poses_.empty()) return;
        const int rows = dims_[0].first, cols = dims_[0].second;
        const size_t P = (size_t)rows * cols;
        out->rows = rows;
        out->cols = cols;
        out->pruned_depth.assign(P, 0.0f);
        out->points3d.resize(3 * P);
        out->colors.resize(3 * P);
        out->keypts.resize(2 * P);
        int64_t n = 0;
        check(ctx_.get(),
              pmp_pruner_prune(ctx_.get(), thr_, out->pruned_depth.data(), (int64_t)P, out->points3d.data(),
                               out->colors.data(), out->keypts.data(), &n),
              "pmp_pruner_prune");
        out->points3d.resize(3 * (size_t)n);
        out->colors.resize(3 * (size_t)n);
        out->keypts.resize(2 * (size_t)n);
        std::memcpy(out->R, poses_[0].R, sizeof(out->R));
        std::memcpy(out->t, poses_[0].t, sizeof(out->t));
    }
    int PopHead() {
        check(ctx_.get(), pmp_pruner_pop_head(ctx_.get()), "pmp_pruner_pop_head");
        if (!poses_.empty()) {
            poses_.erase(poses_.begin());
            dims_.erase(dims_.begin());
        }
        return (int)poses_.size();
    }
    void reset() {
        check(ctx_.get(), pmp_pruner_reset(ctx_.get()), "pmp_pruner_reset");
        poses_.clear();
        dims_.clear();
    }

private:
    Context& ctx_;
    std::vector<pmp_pose> poses_;
    std::vector<std::pair<int, int>> dims_;
    float thr_ = 0.01f;    // depthmap.cc:1379
    int min_views_ = 3;
    int queue_ = 1;
};

}  // namespace pmp
