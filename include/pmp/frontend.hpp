// frontend.hpp -- the densify thread of the OpenVSLAM fork
// (depthmap_computation_module::run, depthmap_computation_module.cc:205-472)
// over the facade classes of densifier.hpp: per incoming keyframe
//   1. parallax gate: against the previous accepted keyframe, the fraction of
//      common landmarks seen under an angle in (pi/60, pi/6) must exceed 0.25
//      (score_view, :147-198; :356-366);
//   2. depth range: [1e-5, 1.1 * 90th percentile of the landmark distances]
//      (compute_depth_intervall, :123-131);
//   3. 3-view estimator window (centre = reference) with the landmark and
//      previous-result pre-initialisation; each depth map goes to
//   4. a 5-deep Cleaner queue (min. 3 consistent views); each cleaned map is
//      returned with the keyframe it belongs to (clean_id, :436-452), ready for
//      the point-cloud thread (DepthmapPruner).
// Header-only C++17 over include/pmp/pmp.h.
#pragma once

#include <cstring>
#include <vector>

#include "pmp/densifier.hpp"

namespace pmp {

// One keyframe as the tracking/mapping side hands it over.
struct FrontendKeyframe {
    uint64_t id = 0;
    uint64_t src_frm_id = 0;
    double R[9];                          // world->camera (rot_cw), row-major
    double t[3];                          // trans_cw
    int width = 0, height = 0;            // thumbnail size
    const uint8_t* gray = nullptr;        // grey thumbnail
    const uint8_t* mask = nullptr;        // mask thumbnail (0 = ignored)
    std::vector<uint64_t> landmark_ids;   // valid landmarks of the keyframe
    std::vector<double> landmark_xyz;     // their world positions, 3 per landmark
};

inline double ScoreView(const FrontendKeyframe& prev, const FrontendKeyframe& cur) {
    double pp[12], pc[12];
    std::memcpy(pp, prev.R, sizeof(prev.R));
    std::memcpy(pp + 9, prev.t, sizeof(prev.t));
    std::memcpy(pc, cur.R, sizeof(cur.R));
    std::memcpy(pc + 9, cur.t, sizeof(cur.t));
    return pmp_score_view(pp, (int64_t)prev.landmark_ids.size(), prev.landmark_ids.data(), prev.landmark_xyz.data(),
                          pc, (int64_t)cur.landmark_ids.size(), cur.landmark_ids.data(), cur.landmark_xyz.data());
}

// false when the keyframe has no landmarks (the reference indexes an empty vector then)
inline bool DepthInterval(const FrontendKeyframe& kf, double& min_depth, double& max_depth) {
    double p[12];
    std::memcpy(p, kf.R, sizeof(kf.R));
    std::memcpy(p + 9, kf.t, sizeof(kf.t));
    return pmp_depth_interval(p, (int64_t)(kf.landmark_xyz.size() / 3), kf.landmark_xyz.data(), &min_depth,
                              &max_depth) == PMP_OK;
}

struct FrontendResult {
    uint64_t keyframe_id = 0;
    DepthmapCleanerResult cleaned;
};

class DensifyFrontend {
public:
    // The module's settings (:211-223; config.cc:150-190 defaults).
    struct Config {
        int neighbours = 2;
        int patch_size = 7;
        int iterations = 2;
        float min_patch_sd = 0.0f;
        double min_score = 0.1;
        float same_depth_threshold = 0.03f;
        int clean_queue = 5;
        int min_consistent_views = 3;
        double parallax_gate = 0.25;
        double fallback_max_depth = 10.0;   // for a keyframe without landmarks
        float clean_max_reproj_px = 0.0f;   // pmp_clean_ex round trip (not in the reference; 0: off)
        float clean_max_normal_deg = 0.0f;  // pmp_clean_ex normal angle (not in the reference; 0: off)
    };

    DensifyFrontend(Context& ctx, const Config& cfg) : cfg_(cfg), est_(ctx), cleaner_(ctx) {
        est_.SetMinImagesCompute(cfg.neighbours + 1);
        est_.SetMinPatchSD(cfg.min_patch_sd);
        est_.SetPatchSize(cfg.patch_size);
        est_.SetPatchMatchIterations(cfg.iterations);
        est_.SetMinScore(cfg.min_score);
        cleaner_.SetMinConsistentViews(cfg.min_consistent_views);
        cleaner_.SetDepthQueueSize(cfg.clean_queue);
        cleaner_.SetSameDepthThreshold(cfg.same_depth_threshold);
        cleaner_.SetMaxReprojectionPx(cfg.clean_max_reproj_px);
        cleaner_.SetMaxNormalAngleDeg(cfg.clean_max_normal_deg);
        reset_ids();
    }

    DepthmapEstimator& estimator() { return est_; }

    // Runs one keyframe through the loop; returns false when the parallax
    // gate rejects it.  Cleaned maps that become ready are appended to out.
    bool AddKeyframe(const FrontendKeyframe& kf, std::vector<FrontendResult>* out) {
        if (have_prev_) {
            const double score = ScoreView(prev_, kf);
            if (!(score > cfg_.parallax_gate)) return false;   // NaN (no common landmark) rejects too
        }
        window_.push_back(kf.id);
        double mind = 1e-5, maxd = cfg_.fallback_max_depth;
        DepthInterval(kf, mind, maxd);
        est_.AddView(kf.R, kf.t, kf.gray, kf.mask, kf.width, kf.height, kf.landmark_xyz, maxd, mind);
        prev_ = kf;
        prev_.gray = prev_.mask = nullptr;
        have_prev_ = true;
        if (est_.IsReadyForCompute()) {
            DepthmapEstimatorResult r;
            est_.ComputePatchMatch(&r);
            ++depth_id_;
            cleaner_.AddView(r);
            est_.PopHeadUnlessOne();
        }
        if (cleaner_.IsReadyForCompute()) {
            FrontendResult res;
            cleaner_.Clean(&res.cleaned);
            res.keyframe_id = window_[clean_id_];
            if (out) out->push_back(std::move(res));
            cleaner_.PopHeadUnlessOne();
            ++clean_id_;
        }
        return true;
    }

    // reset request (:286-314)
    void Reset() {
        est_.reset();
        cleaner_.reset();
        window_.clear();
        have_prev_ = false;
        reset_ids();
    }

private:
    void reset_ids() {
        depth_id_ = cfg_.neighbours / 2;
        clean_id_ = depth_id_ + 2;
    }
    Config cfg_;
    DepthmapEstimator est_;
    DepthmapCleaner cleaner_;
    std::vector<uint64_t> window_;
    FrontendKeyframe prev_;
    bool have_prev_ = false;
    int depth_id_ = 0, clean_id_ = 0;
};

}  // namespace pmp
