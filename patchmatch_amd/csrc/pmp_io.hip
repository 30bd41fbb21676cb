// pmp_io.hip -- output formats of the dense map (host code).
//
//   pmp_write_ply            ASCII PLY "x y z r g b" (map_database_io::
//                            save_point_cloud_dense, io/map_database_io.cc:46-81)
//   pmp_write_dense_msgpack  MessagePack of {"cameras", "keyframes",
//                            "landmarks"} (save_message_pack_dense,
//                            io/map_database_io.cc:22-44), the document the
//                            reference's export_dense_msg_to_ply.py reads
//
// The reference builds an nlohmann::json object and serialises it; objects
// are std::maps, so keys appear in byte order, and ids are decimal strings.
// The writer below emits the same keys in the same order: landmarks numbered
// in keyframe order (dense_pointcloud_module.cc:153-174), one per pruned
// point, with 1st/ref keyframe = the keyframe, n_vis = n_fnd = 1 and the colour
// as (r, g, b) from the pruner's BGR (:166-169); keyframes with the dense
// keypoints as {"pt": [col, row]} (data/common.cc:38-44) and their landmark
// ids.  Sparse-SLAM fields this path does not produce (ORB keypoints,
// descriptors, graph edges) are written empty.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "pmp/pmp.h"

namespace {

struct MsgPack {
    std::vector<uint8_t> b;
    void u8(uint8_t v) { b.push_back(v); }
    void be(uint64_t v, int bytes) {
        for (int k = bytes - 1; k >= 0; --k) b.push_back((uint8_t)(v >> (8 * k)));
    }
    void map(size_t n) {
        if (n < 16) u8((uint8_t)(0x80 | n));
        else if (n < 65536) { u8(0xde); be(n, 2); }
        else { u8(0xdf); be(n, 4); }
    }
    void arr(size_t n) {
        if (n < 16) u8((uint8_t)(0x90 | n));
        else if (n < 65536) { u8(0xdc); be(n, 2); }
        else { u8(0xdd); be(n, 4); }
    }
    void str(const std::string& s) {
        const size_t n = s.size();
        if (n < 32) u8((uint8_t)(0xa0 | n));
        else if (n < 256) { u8(0xd9); be(n, 1); }
        else if (n < 65536) { u8(0xda); be(n, 2); }
        else { u8(0xdb); be(n, 4); }
        b.insert(b.end(), s.begin(), s.end());
    }
    void uint(uint64_t v) {
        if (v < 128) u8((uint8_t)v);
        else if (v < 256) { u8(0xcc); be(v, 1); }
        else if (v < 65536) { u8(0xcd); be(v, 2); }
        else if (v < 4294967296ull) { u8(0xce); be(v, 4); }
        else { u8(0xcf); be(v, 8); }
    }
    void sint(int64_t v) {
        if (v >= 0) { uint((uint64_t)v); return; }
        if (v >= -32) u8((uint8_t)(int8_t)v);
        else if (v >= -128) { u8(0xd0); be((uint8_t)(int8_t)v, 1); }
        else if (v >= -32768) { u8(0xd1); be((uint16_t)(int16_t)v, 2); }
        else if (v >= -2147483648ll) { u8(0xd2); be((uint32_t)(int32_t)v, 4); }
        else { u8(0xd3); be((uint64_t)v, 8); }
    }
    void f64(double d) {
        uint64_t u;
        memcpy(&u, &d, 8);
        u8(0xcb);
        be(u, 8);
    }
};

// Rotation matrix -> unit quaternion (x, y, z, w), w >= 0 branch first
// (Shepperd's method: the largest of w, x, y, z is computed from the trace
// or a diagonal element, the others from off-diagonal sums/differences).
void quat_of(const double* R, double q[4]) {
    const double m00 = R[0], m01 = R[1], m02 = R[2], m10 = R[3], m11 = R[4], m12 = R[5], m20 = R[6], m21 = R[7],
                 m22 = R[8];
    const double t = m00 + m11 + m22;
    if (t > 0.0) {
        double s = sqrt(t + 1.0);
        q[3] = 0.5 * s;
        s = 0.5 / s;
        q[0] = (m21 - m12) * s;
        q[1] = (m02 - m20) * s;
        q[2] = (m10 - m01) * s;
        return;
    }
    int i = 0;
    if (m11 > m00) i = 1;
    if (m22 > (i == 0 ? m00 : m11)) i = 2;
    const int j = (i + 1) % 3, k = (j + 1) % 3;
    auto m = [&](int r, int c) { return R[3 * r + c]; };
    double s = sqrt(m(i, i) - m(j, j) - m(k, k) + 1.0);
    q[i] = 0.5 * s;
    s = 0.5 / s;
    q[3] = (m(k, j) - m(j, k)) * s;
    q[j] = (m(j, i) + m(i, j)) * s;
    q[k] = (m(k, i) + m(i, k)) * s;
}

bool id_less(uint64_t a, uint64_t b) { return std::to_string(a) < std::to_string(b); }

}  // namespace

extern "C" {

int pmp_write_ply(const char* path, int64_t n, const double* xyz, const uint8_t* rgb) {
    if (!path || n < 0 || (n > 0 && (!xyz || !rgb))) return PMP_EINVAL;
    FILE* f = fopen(path, "w");
    if (!f) return PMP_EINVAL;
    fprintf(f,
            "ply\nformat ascii 1.0\nelement vertex %lld\nproperty float x\nproperty float y\nproperty float z\n"
            "property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n",
            (long long)n);
    for (int64_t k = 0; k < n; ++k)
        fprintf(f, "%.6f %.6f %.6f %u %u %u\n", xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], (unsigned)rgb[3 * k],
                (unsigned)rgb[3 * k + 1], (unsigned)rgb[3 * k + 2]);
    const int bad = ferror(f);
    return (fclose(f) != 0 || bad) ? PMP_EDEVICE : PMP_OK;
}

int pmp_write_dense_msgpack(const char* path, const pmp_camera_info* cam, int32_t n_keyframes,
                            const pmp_dense_keyframe* kfs) {
    if (!path || !cam || !cam->name || n_keyframes < 0 || (n_keyframes > 0 && !kfs)) return PMP_EINVAL;
    for (int k = 0; k < n_keyframes; ++k)
        if (kfs[k].n_points < 0 || (kfs[k].n_points > 0 && (!kfs[k].points || !kfs[k].colors || !kfs[k].keypts)))
            return PMP_EINVAL;
    MsgPack m;
    m.map(3);
    // cameras (camera/equirectangular.cc:93-103; keys in byte order)
    m.str("cameras");
    m.map(1);
    m.str(cam->name);
    m.map(9);
    m.str("color_order"); m.str("RGB");
    m.str("cols"); m.uint((uint64_t)cam->cols);
    m.str("focal_x_baseline"); m.f64(0.0);
    m.str("fps"); m.f64(cam->fps);
    m.str("model_type"); m.str("Equirectangular");
    m.str("num_grid_cols"); m.uint(0);
    m.str("num_grid_rows"); m.uint(0);
    m.str("rows"); m.uint((uint64_t)cam->rows);
    m.str("setup_type"); m.str("Monocular");

    // landmark ids: consecutive over the keyframes in the given order
    std::vector<uint64_t> first(n_keyframes + 1, 0);
    for (int k = 0; k < n_keyframes; ++k) first[k + 1] = first[k] + (uint64_t)kfs[k].n_points;

    // keyframes (data/keyframe.cc:138-200 with bDense)
    std::vector<int> order(n_keyframes);
    for (int k = 0; k < n_keyframes; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return id_less(kfs[a].id, kfs[b].id); });
    m.str("keyframes");
    m.map((size_t)n_keyframes);
    for (int o : order) {
        const pmp_dense_keyframe& kf = kfs[o];
        m.str(std::to_string(kf.id));
        m.map(18);
        m.str("cam"); m.str(cam->name);
        m.str("depth_thr"); m.f64(0.0);
        m.str("depths"); m.arr(0);
        m.str("descs"); m.arr(0);
        m.str("keypts");
        m.arr((size_t)kf.n_points);
        for (int64_t q = 0; q < kf.n_points; ++q) {   // {"pt": [col, row]} (data/common.cc:38-44)
            m.map(1);
            m.str("pt");
            m.arr(2);
            m.sint(kf.keypts[2 * q + 1]);
            m.sint(kf.keypts[2 * q]);
        }
        m.str("lm_ids");
        m.arr((size_t)kf.n_points);
        for (int64_t q = 0; q < kf.n_points; ++q) m.uint(first[o] + (uint64_t)q);
        m.str("loop_edges"); m.arr(0);
        m.str("n_keypts"); m.uint(0);
        m.str("n_scale_levels"); m.uint(0);
        double qv[4];
        quat_of(kf.pose.R, qv);
        m.str("rot_cw");
        m.arr(4);
        for (double v : qv) m.f64(v);
        m.str("scale_factor"); m.f64(1.0);
        m.str("span_children"); m.arr(0);
        m.str("span_parent"); m.sint(-1);
        m.str("src_frm_id"); m.uint(kf.src_frm_id);
        m.str("trans_cw");
        m.arr(3);
        for (int a = 0; a < 3; ++a) m.f64(kf.pose.t[a]);
        m.str("ts"); m.f64(kf.timestamp);
        m.str("undists"); m.arr(0);
        m.str("x_rights"); m.arr(0);
    }

    // landmarks (data/landmark.cc:370-377), ids as decimal strings in byte order
    std::vector<uint64_t> ids(first[n_keyframes]);
    for (uint64_t q = 0; q < ids.size(); ++q) ids[q] = q;
    std::sort(ids.begin(), ids.end(), id_less);
    m.str("landmarks");
    m.map(ids.size());
    int kf_of = 0;
    for (uint64_t id : ids) {
        kf_of = (int)(std::upper_bound(first.begin(), first.end(), id) - first.begin()) - 1;
        const pmp_dense_keyframe& kf = kfs[kf_of];
        const uint64_t q = id - first[kf_of];
        m.str(std::to_string(id));
        m.map(6);
        m.str("1st_keyfrm"); m.uint(kf.id);
        m.str("color");                       // setColor(c_r, c_g, c_b) from BGR
        m.arr(3);
        m.f64((double)kf.colors[3 * q + 2]);
        m.f64((double)kf.colors[3 * q + 1]);
        m.f64((double)kf.colors[3 * q]);
        m.str("n_fnd"); m.uint(1);
        m.str("n_vis"); m.uint(1);
        m.str("pos_w");
        m.arr(3);
        for (int a = 0; a < 3; ++a) m.f64(kf.points[3 * q + a]);
        m.str("ref_keyfrm"); m.uint(kf.id);
    }
    FILE* f = fopen(path, "wb");
    if (!f) return PMP_EINVAL;
    const size_t w = fwrite(m.b.data(), 1, m.b.size(), f);
    return (fclose(f) != 0 || w != m.b.size()) ? PMP_EDEVICE : PMP_OK;
}

}  // extern "C"
