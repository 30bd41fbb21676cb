// pmp_frontend.hip -- the keyframe gate and depth range of the densify thread
// (host code): depthmap_computation_module::score_view (:147-198) and
// compute_depth_intervall (:123-131) with keyframe::computePercentileDepth_10_90
// (data/keyframe.cc:428-451).  include/pmp/frontend.hpp builds the densify
// thread's loop on them.
#include <math.h>

#include <algorithm>
#include <unordered_map>
#include <vector>

#include "pmp/pmp.h"

extern "C" {

double pmp_score_view(const double* prev_pose, int64_t n_prev, const uint64_t* prev_ids, const double* prev_xyz,
                      const double* cur_pose, int64_t n_cur, const uint64_t* cur_ids, const double* cur_xyz) {
    (void)prev_xyz;
    if (!prev_pose || !cur_pose || n_prev < 0 || n_cur < 0 || (n_prev && !prev_ids) ||
        (n_cur && (!cur_ids || !cur_xyz)))
        return NAN;
    // camera centres -R^T t
    double cp[3], cc[3];
    for (int r = 0; r < 3; ++r) {
        cp[r] = -(prev_pose[0 + r] * prev_pose[9] + prev_pose[3 + r] * prev_pose[10] + prev_pose[6 + r] * prev_pose[11]);
        cc[r] = -(cur_pose[0 + r] * cur_pose[9] + cur_pose[3 + r] * cur_pose[10] + cur_pose[6 + r] * cur_pose[11]);
    }
    std::unordered_map<uint64_t, int64_t> in_prev;
    in_prev.reserve((size_t)n_prev * 2);
    for (int64_t k = 0; k < n_prev; ++k) in_prev.emplace(prev_ids[k], k);
    const double theta_min = M_PI / 60.0, theta_max = M_PI / 6.0;
    double score = 0.0;
    int64_t common = 0;
    for (int64_t k = 0; k < n_cur; ++k) {
        if (!in_prev.count(cur_ids[k])) continue;
        ++common;
        const double* lm = cur_xyz + 3 * k;
        double dp[3], dc[3];
        for (int a = 0; a < 3; ++a) {
            dp[a] = cp[a] - lm[a];
            dc[a] = cc[a] - lm[a];
        }
        const double nom = dc[0] * dp[0] + dc[1] * dp[1] + dc[2] * dp[2];
        const double v1 = dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2];
        const double v2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
        const double theta = acos(nom / sqrt(v1 * v2));
        if (theta > theta_min && theta < theta_max) score += 1.0;
    }
    return score / (double)common;   // NaN without common landmarks, as the reference's 0/0
}

int pmp_depth_interval(const double* pose, int64_t n, const double* xyz, double* min_depth, double* max_depth) {
    if (!pose || !min_depth || !max_depth || n < 0 || (n > 0 && !xyz)) return PMP_EINVAL;
    if (n == 0) return PMP_EINVAL;   // the reference indexes an empty vector here
    std::vector<float> d((size_t)n);
    for (int64_t k = 0; k < n; ++k) {
        const double* X = xyz + 3 * k;
        double c[3];
        for (int r = 0; r < 3; ++r) c[r] = pose[3 * r] * X[0] + pose[3 * r + 1] * X[1] + pose[3 * r + 2] * X[2] + pose[9 + r];
        d[k] = (float)sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    }
    std::sort(d.begin(), d.end());
    const double p90 = d[(size_t)n / 10 * 9];   // index size/10*9 (data/keyframe.cc:450)
    *min_depth = 1e-5;                           // depthmap_computation_module.cc:129
    *max_depth = p90 * 1.1;                      // :130
    return PMP_OK;
}

}  // extern "C"
