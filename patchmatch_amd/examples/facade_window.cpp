// facade_window.cpp -- drives pmp::DepthmapEstimator (include/pmp/densifier.hpp)
// the way the OpenVSLAM densify thread drives dense::DepthmapEstimator
// (depthmap_computation_module.cc:374-397): AddView per keyframe, and once the
// window is full, ComputePatchMatch (with the landmark and previous-result
// pre-initialisation) + PopHeadUnlessOne.
//
//   facade_window <input.bin> <output.bin>
// input : int32 W, H, n_views; pmp_params (raw struct); per view: 12 doubles
//         (R row-major, t), W*H grey bytes, W*H mask bytes, double mind, maxd,
//         int64 n_landmarks, n_landmarks*3 doubles (world xyz)
// output: per computed window: W*H float depth, W*H float score, W*H int32 nghbr
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pmp/densifier.hpp"

static bool rd(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    FILE* out = std::fopen(argv[2], "wb");
    if (!in || !out) return 2;
    int32_t hdr[3];
    pmp_params prm;
    if (!rd(in, hdr, sizeof(hdr)) || !rd(in, &prm, sizeof(prm))) return 2;
    const int W = hdr[0], H = hdr[1], n = hdr[2];
    try {
        pmp::Context ctx(0);
        pmp::DepthmapEstimator est(ctx);
        est.params() = prm;
        est.SetMinImagesCompute(3);
        for (int v = 0; v < n; ++v) {
            double pose[12], mind, maxd;
            std::vector<uint8_t> img((size_t)W * H), mask((size_t)W * H);
            int64_t nl = 0;
            if (!rd(in, pose, sizeof(pose)) || !rd(in, img.data(), img.size()) || !rd(in, mask.data(), mask.size()) ||
                !rd(in, &mind, 8) || !rd(in, &maxd, 8) || !rd(in, &nl, 8) || nl < 0)
                return 2;
            std::vector<double> lms((size_t)nl * 3);
            if (nl > 0 && !rd(in, lms.data(), lms.size() * sizeof(double))) return 2;
            est.AddView(pose, pose + 9, img.data(), mask.data(), W, H, lms, maxd, mind);
            if (est.IsReadyForCompute()) {
                pmp::DepthmapEstimatorResult r;
                est.ComputePatchMatch(&r);
                std::fwrite(r.depth.data(), sizeof(float), r.depth.size(), out);
                std::fwrite(r.score.data(), sizeof(float), r.score.size(), out);
                std::fwrite(r.nghbr.data(), sizeof(int32_t), r.nghbr.size(), out);
                est.PopHeadUnlessOne();
            }
        }
    } catch (const pmp::Error& e) {
        std::fprintf(stderr, "pmp error %d: %s\n", e.code, e.what());
        return 1;
    }
    std::fclose(in);
    std::fclose(out);
    return 0;
}
