// frontend_stream.cpp -- runs a keyframe stream through pmp::DensifyFrontend
// (include/pmp/frontend.hpp), the densify thread's loop
// (depthmap_computation_module.cc:205-472): gate, depth range, 3-view window
// with pre-initialisation, 5-deep Cleaner.
//
//   frontend_stream <input.bin> <output.bin>
// input : int32 W, H, n, iterations; float same_depth_threshold,
//         clean_max_reproj_px, clean_max_normal_deg (0: off); pmp_params
//         (raw struct, the estimator's base params); per keyframe: uint64 id,
//         12 doubles (R row-major, t), W*H grey, W*H mask, int64 n_landmarks,
//         n_landmarks uint64 ids, n_landmarks*3 doubles (world xyz)
// output: per keyframe: int32 accepted; then per cleaned result: uint64
//         keyframe id, W*H float cleaned depth (in the order they appear)
#include <cstdio>
#include <vector>

#include "pmp/frontend.hpp"

static bool rd(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    FILE* out = std::fopen(argv[2], "wb");
    if (!in || !out) return 2;
    int32_t hdr[4];
    float thr[3];
    pmp_params prm;
    if (!rd(in, hdr, sizeof(hdr)) || !rd(in, thr, sizeof(thr)) || !rd(in, &prm, sizeof(prm))) return 2;
    const int W = hdr[0], H = hdr[1], n = hdr[2];
    try {
        pmp::Context ctx(0);
        pmp::DensifyFrontend::Config cfg;
        cfg.iterations = hdr[3];
        cfg.same_depth_threshold = thr[0];
        cfg.clean_max_reproj_px = thr[1];
        cfg.clean_max_normal_deg = thr[2];
        pmp::DensifyFrontend fe(ctx, cfg);
        pmp_params& p = fe.estimator().params();
        const int iters = p.iterations, half = p.patch_half;
        p = prm;
        p.iterations = iters;
        p.patch_half = half;
        std::vector<pmp::FrontendResult> results;
        for (int k = 0; k < n; ++k) {
            pmp::FrontendKeyframe kf;
            double pose[12];
            std::vector<uint8_t> img((size_t)W * H), mask((size_t)W * H);
            int64_t nl = 0;
            if (!rd(in, &kf.id, 8) || !rd(in, pose, sizeof(pose)) || !rd(in, img.data(), img.size()) ||
                !rd(in, mask.data(), mask.size()) || !rd(in, &nl, 8) || nl < 0)
                return 2;
            kf.landmark_ids.resize((size_t)nl);
            kf.landmark_xyz.resize((size_t)nl * 3);
            if (nl > 0 && (!rd(in, kf.landmark_ids.data(), (size_t)nl * 8) ||
                           !rd(in, kf.landmark_xyz.data(), (size_t)nl * 24)))
                return 2;
            for (int a = 0; a < 9; ++a) kf.R[a] = pose[a];
            for (int a = 0; a < 3; ++a) kf.t[a] = pose[9 + a];
            kf.width = W;
            kf.height = H;
            kf.gray = img.data();
            kf.mask = mask.data();
            const int32_t ok = fe.AddKeyframe(kf, &results) ? 1 : 0;
            std::fwrite(&ok, 4, 1, out);
        }
        for (const auto& r : results) {
            std::fwrite(&r.keyframe_id, 8, 1, out);
            std::fwrite(r.cleaned.cleaned_depth.data(), sizeof(float), r.cleaned.cleaned_depth.size(), out);
        }
    } catch (const pmp::Error& e) {
        std::fprintf(stderr, "pmp error %d: %s\n", e.code, e.what());
        return 1;
    }
    std::fclose(in);
    std::fclose(out);
    return 0;
}
