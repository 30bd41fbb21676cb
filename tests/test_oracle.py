"""CPU checks of the oracle: published known-answer vectors, libm error
bounds, closed forms, invariants and the FP64 scorer.  The reference has no
tests or fixtures for this path (its test/ tree is upstream OpenVSLAM), so
each oracle function is pinned here by mathematics chosen so that a dropped
term, a wrong sign or index, or a transposed operand fails."""
import json
import os

import numpy as np
import pytest

from oracle import or_py
from patchmatch_amd import synth
from patchmatch_amd.params import (AGG_FOLD_INDEP, AGG_MAX, AGG_REF_FOLD, AGG_TOPK_MEAN, COVER_FULL, EDGE_WRAP,
                                   config0_params, config1_params, default_params)

from conftest import gt_planes, random_planes

HERE = os.path.dirname(os.path.abspath(__file__))


def ulp_err(approx, exact):
    approx = np.asarray(approx, np.float64)
    exact = np.asarray(exact, np.float64)
    ulp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    return np.abs(approx - exact) / np.maximum(ulp, 1e-45)


# ------------------------------------------------------------------ RNG
def test_philox_known_answer_vectors():
    kat = json.load(open(os.path.join(HERE, "golden", "philox4x32_10_kat.json")))
    for v in kat["vectors"]:
        out = or_py.philox([int(x, 16) for x in v["ctr"]], [int(x, 16) for x in v["key"]])
        assert out == [int(x, 16) for x in v["out"]]


def test_refine_normals_moments(kf_c0):
    """Box-Muller normals of the refinement stream: mean 0, variance 1, no
    correlation between pixels / iterations (counter-based streams)."""
    ok = or_py.Keyframe.from_synth(kf_c0)
    p = config0_params(seed=99)
    z = np.concatenate([or_py.refine_normals(ok, p, i, j, it, 18)
                        for i in range(0, 200, 7) for j in range(0, 400, 11) for it in range(2)])
    assert abs(z.mean()) < 0.02 and abs(z.var() - 1.0) < 0.03
    a = or_py.refine_normals(ok, p, 10, 10, 0, 18)
    b = or_py.refine_normals(ok, p, 10, 11, 0, 18)
    c = or_py.refine_normals(ok, p, 10, 10, 1, 18)
    assert not np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.array_equal(a, or_py.refine_normals(ok, p, 10, 10, 0, 18))


# ------------------------------------------------------------------ elementary functions
@pytest.mark.parametrize("name,ref,dom,bound", [
    ("asin", np.arcsin, (-1, 1), 4.0),
    ("acos", np.arccos, (-1, 1), 3.0),
    ("exp", np.exp, (-80, 80), 1.5),
    ("rsqrt", lambda x: 1 / np.sqrt(x), (1e-6, 1e6), 1.0),
    ("sqrt", np.sqrt, (1e-6, 1e6), 2.0),
    ("rcp", lambda x: 1 / x, (1e-6, 1e6), 0.51),
])
def test_math_vs_libm(name, ref, dom, bound):
    rng = np.random.default_rng(1)
    if name in ("rsqrt", "sqrt", "rcp"):
        x = (10 ** rng.uniform(np.log10(dom[0]), np.log10(dom[1]), 200000)).astype(np.float32)
    else:
        x = rng.uniform(dom[0], dom[1], 200000).astype(np.float32)
    y = or_py.eval_math(name, x)
    e = ulp_err(y, ref(x.astype(np.float64)))
    assert e.max() <= bound, (name, e.max(), x[np.argmax(e)])


def test_atan2_vs_libm():
    rng = np.random.default_rng(2)
    a = rng.normal(0, 5, 200000).astype(np.float32)
    b = rng.normal(0, 5, 200000).astype(np.float32)
    y = or_py.eval_math("atan2", a, b)
    assert ulp_err(y, np.arctan2(a.astype(np.float64), b.astype(np.float64))).max() <= 3.0
    assert or_py.eval_math("atan2", np.float32([0.0]), np.float32([0.0]))[0] == 0.0
    # the sign of a zero y is kept, as C atan2: atan2(-0, -1) = -pi, atan2(-0, 1) = -0
    z = or_py.eval_math("atan2", np.float32([-0.0, -0.0, 0.0]), np.float32([-1.0, 1.0, -1.0]))
    assert z[0] == np.float32(-np.pi) and z[1] == 0.0 and np.signbit(z[1]) and z[2] == np.float32(np.pi)


def test_atan2_ratio_needs_no_clamp():
    """atan2 and the latitude form their ratio in [0, 1] without a clamp:
    m rcp(m) <= 1 for every float m in [1, 2) (the reciprocal's seed and steps
    scale exactly with the binade; monotone rounding covers mn < mx), so the
    latitude's 1/t above 45 degrees (t > 1) is below 1 too."""
    m = (np.arange(1 << 23, dtype=np.uint32) | np.uint32(0x3F800000)).view(np.float32)
    assert (m * or_py.eval_math("rcp", m)).astype(np.float32).max() <= 1.0
    # both poles and the horizon without a clamp on h^2: (b, a, b) with b = 0
    lat = or_py.eval_math("lat", np.float32([2.0, -3.0, 0.0]), np.float32([0.0, 0.0, 0.0]))
    assert lat[0] == np.float32(np.pi / 2) and lat[-2] == -np.float32(np.pi / 2) and lat[2] == 0.0


def test_longitude_vs_libm():
    """Longitude of the direction (a, 0, b): atan2(a, b) from the octant angle
    asin(min(|a|, |b|) / h), every quadrant and both octants of each."""
    rng = np.random.default_rng(4)
    a = np.concatenate([rng.normal(0, 5, 300000), rng.normal(0, 1e-3, 1000), rng.uniform(-1, 1, 1000)])
    b = np.concatenate([rng.normal(0, 5, 300000), rng.normal(0, 5, 1000), rng.uniform(-1e-3, 1e-3, 1000)])
    a, b = a.astype(np.float32), b.astype(np.float32)
    lon = or_py.eval_math("lon", a, b)
    exact = np.arctan2(a.astype(np.float64), b.astype(np.float64))
    assert ulp_err(lon, exact).max() <= 4.5
    assert np.abs(lon - exact).max() <= 4e-7
    # axes and signed zeros as C atan2: (0, 1) -> 0, (0, -1) -> pi, (-0, -1) -> -pi, (+-1, 0) -> +-pi/2
    sp = or_py.eval_math("lon", np.float32([0.0, 0.0, -0.0, 1.0, -1.0, 0.0]), np.float32([1.0, -1.0, -1.0, 0.0, 0.0, 0.0]))
    assert sp[0] == 0.0 and sp[1] == np.float32(np.pi) and sp[2] == -np.float32(np.pi)
    assert sp[3] == np.float32(np.pi / 2) and sp[4] == -np.float32(np.pi / 2) and sp[5] == 0.0
    # the diagonal x = z over two binades (the octant boundary, where asin's
    # slope is sqrt 2 and r = rsqrt(2 m^2) repeats every two binades): pi/4 within 3 ulp
    m = (np.arange(1 << 24, dtype=np.uint32) + np.uint32(0x3F800000)).view(np.float32)
    d = or_py.eval_math("lon", m, m).astype(np.float64)
    assert np.abs(d - np.pi / 4).max() <= 3 * np.spacing(np.float32(np.pi / 4))


def test_latitude_vs_libm():
    """Latitude of the direction (x, y, z) = (b, a, b): atan2(a, sqrt(2) |b|),
    over both branches (|y| below and above the horizontal norm) and the poles."""
    rng = np.random.default_rng(3)
    a = np.concatenate([rng.normal(0, 5, 200000), rng.normal(0, 1e-3, 1000), [1.0, -1.0, 0.0, 3.0]])
    b = np.concatenate([rng.normal(0, 5, 200000), rng.normal(0, 5, 1000), [0.0, 0.0, 1.0, 3.0 / np.sqrt(2)]])
    a, b = a.astype(np.float32), b.astype(np.float32)
    lat = or_py.eval_math("lat", a, b)
    h = np.sqrt(2.0) * np.abs(b.astype(np.float64))
    exact = np.arctan2(a.astype(np.float64), h)
    assert ulp_err(lat, exact).max() <= 4.0
    assert lat[-4] == np.float32(np.pi / 2) and lat[-3] == -np.float32(np.pi / 2) and lat[-2] == 0.0


def test_log_and_sincos():
    u = ((np.arange(1, 1 << 24, 97) + 0.0) * 2.0 ** -24).astype(np.float32)
    lg = np.log(u.astype(np.float64))
    nz = np.abs(lg) > 1e-3
    assert ulp_err(or_py.eval_math("log", u)[nz], lg[nz]).max() <= 2.0
    v = (np.arange(0, 1 << 24, 89) * 2.0 ** -24).astype(np.float32)
    assert np.abs(or_py.eval_math("sin2pi", v) - np.sin(2 * np.pi * v.astype(np.float64))).max() < 3e-7
    assert np.abs(or_py.eval_math("cos2pi", v) - np.cos(2 * np.pi * v.astype(np.float64))).max() < 3e-7


def test_math_special_values():
    assert or_py.eval_math("exp", np.float32([-100.0, np.nan]))[0] == 0.0
    assert np.isfinite(or_py.eval_math("rcp", np.float32([0.0, np.nan, np.inf]))).all()
    assert or_py.eval_math("sqrt", np.float32([0.0, -1.0, np.nan])).tolist() == [0.0, 0.0, 0.0]
    assert np.isclose(or_py.eval_math("acos", np.float32([1.0]))[0], 0.0)
    assert np.isclose(or_py.eval_math("asin", np.float32([1.0]))[0], np.pi / 2, rtol=1e-7)


# ------------------------------------------------------------------ bearings (PatchMatchCuda.cu:191-198)
def test_bearing_tables_closed_form():
    W, H = 480, 240
    rt, ct = or_py.tables(W, H)
    bx = rt[:, 0:1] * ct[None, :, 0]
    by = -np.broadcast_to(rt[:, 1:2], (H, W))
    bz = rt[:, 0:1] * ct[None, :, 1]
    n = np.sqrt(bx.astype(np.float64) ** 2 + by.astype(np.float64) ** 2 + bz.astype(np.float64) ** 2)
    assert np.abs(n - 1).max() < 3e-7
    assert (bx[H // 2, W // 2], by[H // 2, W // 2], bz[H // 2, W // 2]) == (0.0, 0.0, 1.0)
    assert by[0, 0] == -1.0 and abs(bz[0, 0]) < 1e-7           # top row: straight up (y down)
    assert abs(bx[H // 2, 0]) < 1e-7 and bz[H // 2, 0] == -1.0   # column 0: lon = -pi, looking backwards
    assert bx[H // 2, W // 4] == -1.0                             # lon = -pi/2: -x
    ref = synth.bearings(W, H)
    assert np.abs(np.stack([bx, by, bz], -1) - ref).max() < 2e-7


# ------------------------------------------------------------------ warp (PatchMatchCuda.cu:111-146)
def _kf_pair(ref, view, ref_pose, view_pose):
    return or_py.Keyframe(ref, [view], ref_pose, np.asarray(view_pose)[None])


def test_identity_pose_maps_tap_to_itself():
    img = np.zeros((120, 240), np.uint8)
    pose = synth.yaw_pose(0.3, [0.1, 0.0, -0.2])
    kf = _kf_pair(img, img, pose, pose)
    p = default_params()
    rng = np.random.default_rng(0)
    for _ in range(50):
        i, j = int(rng.integers(5, 115)), int(rng.integers(0, 240))
        d = rng.uniform(0.5, 8.0)
        n = rng.normal(size=3)
        b = synth.bearings(240, 120)[i, j]
        n = n if n @ b < 0 else -n
        plane = n / (n @ (d * b))
        u, v = or_py.warp_point(kf, p, i, j, 0, plane)
        du = (u - j + 120) % 240 - 120
        assert abs(du) < 2e-4 and abs(v - i) < 2e-4, (i, j, u, v)


def test_pure_rotation_is_plane_independent():
    img = np.zeros((120, 240), np.uint8)
    c = [0.2, -0.1, 0.3]
    ref_pose = synth.yaw_pose(0.0, c)
    view_pose = synth.yaw_pose(2 * np.pi * 17 / 240, c)       # 17-column yaw, same centre
    kf = _kf_pair(img, img, ref_pose, view_pose)
    p = default_params()
    for (i, j) in [(30, 40), (60, 200), (100, 5)]:
        b = synth.bearings(240, 120)[i, j]
        uvs = []
        for d, n in [(1.0, -b), (5.0, -b + np.array([0.3, 0.1, 0.0])), (0.4, np.array([0.0, 1.0, 0.2]))]:
            n = n if n @ b < 0 else -n
            uvs.append(or_py.warp_point(kf, p, i, j, 0, n / (n @ (d * b))))
        uvs = np.array(uvs)
        assert np.ptp(uvs[:, 0]) < 2e-4 and np.ptp(uvs[:, 1]) < 2e-4
        # rotating the camera by +yaw moves the scene by -17 columns
        assert abs(((uvs[0, 0] - (j - 17)) + 120) % 240 - 120) < 2e-4 and abs(uvs[0, 1] - i) < 2e-4


def test_translated_camera_wall_analytic():
    """Plane z = 3 in front of the reference, neighbour translated by (0.4, -0.1, 0.25):
    (u, v) against the double-precision closed form, within 1e-3 px."""
    W, H = 960, 480
    img = np.zeros((H, W), np.uint8)
    ref_pose = synth.yaw_pose(0.0, [0, 0, 0])
    t = np.array([0.4, -0.1, 0.25])
    view_pose = synth.yaw_pose(0.0, t)
    kf = _kf_pair(img, img, ref_pose, view_pose)
    p = default_params()
    plane = np.array([0.0, 0.0, 1.0 / 3.0])
    B = synth.bearings(W, H)
    worst = 0.0
    for i in range(140, 340, 23):
        for j in range(300, 660, 37):
            b = B[i, j]
            X = b / (plane @ b)
            Xn = X - t
            u = W * (0.5 + np.arctan2(Xn[0], Xn[2]) / (2 * np.pi))
            v = H * (0.5 + np.arcsin(Xn[1] / np.linalg.norm(Xn)) / np.pi)
            uu, vv = or_py.warp_point(kf, p, i, j, 0, plane)
            worst = max(worst, abs(uu - u), abs(vv - v))
    assert worst < 1e-3, worst


# ------------------------------------------------------------------ NCC (PatchMatchCuda.cu:148-171)
def _textured(H=96, W=192, seed=0):
    return synth.render(synth.make_scene(seed), synth.yaw_pose(0.0, [0, 0, 0]), W, H)[0].round().astype(np.uint8)


def _score_one_view(ref, view, plane_fn, p, pose=None):
    pose = synth.yaw_pose(0.0, [0, 0, 0]) if pose is None else pose
    kf = _kf_pair(ref, view, pose, pose)
    H, W = ref.shape
    p4 = np.zeros((H, W, 4), np.float32)
    B = synth.bearings(W, H)
    p4[..., :3] = plane_fn(B)
    return or_py.score_field(kf, p, p4, views=True)[1][..., 0]


def test_ncc_identity_is_one_and_affine_invariant():
    ref = _textured()
    p = default_params(patch_half=3, patch_stride=1, agg_mode=AGG_MAX)
    fronto = lambda B: -B / -2.0
    s1 = _score_one_view(ref, ref, fronto, p)
    s_shift = _score_one_view(ref, np.clip(ref.astype(int) + 9, 0, 255).astype(np.uint8), fronto, p)
    s_neg = _score_one_view(ref, (255 - ref).astype(np.uint8), fronto, p)
    ok = np.isfinite(s1) & (s1 > -1)
    assert ok.mean() > 0.85          # interior of 96x192 with half 3 is 89 %
    # FP32 warp coordinates carry ~1e-4 px of rounding: NCC within 1e-5 of the
    # exact value on 99 % of pixels, 1e-4 everywhere
    from scipy.ndimage import maximum_filter
    unclipped = maximum_filter(ref, size=9) < 246          # the +9 shift clips nowhere in the patch
    for s, target, m in ((s1, 1.0, ok), (s_shift, 1.0, ok & unclipped), (s_neg, -1.0, ok)):
        d = np.abs(s[m] - target)
        assert np.mean(d <= 1e-5) >= 0.99 and d.max() < 1e-4


def test_ncc_variance_floor():
    ref = np.full((64, 128), 77, np.uint8)
    p = default_params(patch_half=3, patch_stride=1)
    s = _score_one_view(ref, _textured(64, 128), lambda B: B * 0.5, p)
    assert (s[np.isfinite(s)] == -1).all()


@pytest.mark.parametrize("agg", [AGG_REF_FOLD, AGG_FOLD_INDEP])
def test_view_aggregation_truth_tables(agg):
    # the fold itself (PatchMatchCuda.cu:171); REF_FOLD applies it to pooled per-view scores
    fold = default_params(agg_mode=agg)
    # s=-1, s2=0.8 -> |(-1-0.8)| < 0.25*0.8 false -> max -> 0.8; then s2=0.7: |0.1| < 0.175 -> avg 0.75
    assert np.isclose(or_py.aggregate(fold, [0.8, 0.7]), 0.75)
    assert np.isclose(or_py.aggregate(fold, [0.8, 0.3]), 0.8)          # no agreement -> max
    assert or_py.aggregate(fold, [0.0, 0.0]) == 0.0                    # s_v = 0: 0/0 path -> max
    assert or_py.aggregate(fold, [-1.0, -1.0]) == -1.0
    assert np.isclose(or_py.aggregate(fold, [0.5, 0.0]), 0.5)          # division by 0 -> max
    assert or_py.aggregate(default_params(agg_mode=AGG_MAX), [0.1, 0.9, -0.3]) == np.float32(0.9)
    top = default_params(agg_mode=AGG_TOPK_MEAN, topk=3)
    assert np.isclose(or_py.aggregate(top, [0.1, 0.9, -0.3, 0.5, 0.7, 0.2]), (0.9 + 0.7 + 0.5) / 3)
    assert np.isclose(or_py.aggregate(default_params(agg_mode=AGG_TOPK_MEAN, topk=9), [0.5, 0.1]), 0.3)


def _weighted_ref_stats(ref, half, stride, sigma_c, sigma_s):
    """Bilateral-weighted mean and variance of every pixel's reference taps, in
    float64 from the weight's definition (PatchMatchCuda.cu:149): taps at
    offsets -half..half by stride, w = exp(-(x - x_centre)^2 / (2 sc^2) -
    |k|^2 / (2 ss^2)).  NaN where the patch leaves the image."""
    H, W = ref.shape
    x0 = ref.astype(np.float64)
    sw = np.zeros((H, W)); swx = np.zeros((H, W)); swxx = np.zeros((H, W))
    ks = range(-half, half + 1, stride)
    for k1 in ks:
        for k2 in ks:
            x = np.roll(np.roll(x0, -k1, 0), -k2, 1)
            w = np.exp(-(x - x0) ** 2 / (2 * sigma_c ** 2) - (k1 * k1 + k2 * k2) / (2 * sigma_s ** 2))
            sw += w; swx += w * x; swxx += w * x * x
    mx = swx / sw
    vx = swxx / sw - mx * mx
    edge = np.zeros((H, W), bool)
    edge[:half + 1] = edge[H - half - 1:] = True
    edge[:, :half + 1] = edge[:, W - half - 1:] = True
    mx[edge] = np.nan
    vx[edge] = np.nan
    return mx, vx


def test_pooled_fold_two_view_closed_form():
    """REF_FOLD pools the moment sums across views, as PatchMatchCuda.cu does
    (zeroed once at :57-62, outside the view loop :101; NCC of the running
    sums at :161-169).  Identity poses and views that are exact affine maps of
    the reference, y0 = x + 30 and y1 = 2x + 10, in integers: view 0's NCC is 1;
    view 1's is that of the two views' samples together,
        cov = (vx + 2 vx)/2,  var_y = (vx + 4 vx)/2 + ((m0 - m1)/2)^2,
    m0 - m1 = (mx + 30) - (2 mx + 10) = 20 - mx, with mx, vx the weighted mean
    and variance of the reference taps.  Independent per-view NCCs
    (FOLD_INDEP, MAX) are 1 and 1.  The FP32 oracle (per-view shifted moments
    pooled by running sums and Welford's update) and the FP64 scorer (raw
    sums that run on across views) must both give the closed form."""
    base = _textured(96, 192, seed=3)
    ref = (base.astype(np.int32) * 100 // 255).astype(np.uint8)
    views = [(ref.astype(np.int32) + 30).astype(np.uint8), (2 * ref.astype(np.int32) + 10).astype(np.uint8)]
    pose = synth.yaw_pose(0.0, [0, 0, 0])
    kf = or_py.Keyframe(ref, views, pose, np.stack([pose, pose]))
    H, W = ref.shape
    B = synth.bearings(W, H)
    p4 = np.zeros((H, W, 4), np.float32)
    p4[..., :3] = B * 0.5                  # fronto-parallel, depth 2 (any plane: identity poses)
    half, stride = 3, 1
    p = default_params(patch_half=half, patch_stride=stride, agg_mode=AGG_REF_FOLD)
    mx, vx = _weighted_ref_stats(ref, half, stride, p.sigma_color, p.sigma_space)
    e1 = 1.5 * vx / np.sqrt(vx * (2.5 * vx + ((20.0 - mx) / 2.0) ** 2))
    fold = np.where(np.abs(1.0 - e1) < 0.25 * e1, 0.5 * (1.0 + e1), np.maximum(1.0, e1))
    m = np.isfinite(vx) & (vx > 1.0)
    assert m.mean() > 0.8
    for name, fn in (("fp32", or_py.score_field), ("fp64", or_py.score_field_f64)):
        s, sv = fn(kf, p, p4, views=True)
        d0 = np.abs(sv[..., 0][m] - 1.0)
        d1 = np.abs(sv[..., 1][m] - e1[m])
        ds = np.abs(s[m] - fold[m])
        print(f"{name}: view0 max {d0.max():.1e}, pooled view1 max {d1.max():.1e}, folded max {ds.max():.1e}")
        # identity-pose warps carry ~1e-4 px of FP32 rounding (FP64: none)
        assert np.mean(d1 <= 1e-5) >= 0.99 and d1.max() < 1e-3, name
        assert np.mean(d0 <= 1e-5) >= 0.99 and ds.max() < 1e-3, name
    # the pooled value is not the independent one: it drops below 1 by the
    # spread of the view means, here by more than 0.01 on most pixels
    assert np.mean(1.0 - e1[m] > 0.01) > 0.5
    for agg in (AGG_FOLD_INDEP, AGG_MAX):
        s, sv = or_py.score_field(kf, p.copy(agg_mode=agg), p4, views=True)
        assert np.mean(np.abs(sv[..., 1][m] - 1.0) <= 1e-5) >= 0.99
        assert np.mean(np.abs(s[m] - 1.0) <= 1e-5) >= 0.99


# ------------------------------------------------------------------ FP32 vs FP64 (fixed hypothesis field)
@pytest.mark.parametrize("cfg", ["c0", "c0_topk", "small6"])
def test_fp32_score_vs_fp64_oracle(cfg, kf_c0, kf_small6):
    """The NCC of a fixed hypothesis field against the FP64 scorer (exact
    bearings, two-step transform, libm functions, single-exp weight, raw
    moments; for c0's REF_FOLD the raw sums run on across views, as in the
    reference), warp and NCC errors together.  Per view: |d| <= 1e-5 on
    >= 99.9 % of interior pixels and <= 2e-3 everywhere; aggregated: <= 1e-5
    on >= 99.9 %."""
    kf = kf_small6 if cfg == "small6" else kf_c0
    p = {"c0": config0_params(), "c0_topk": config0_params(agg_mode=AGG_TOPK_MEAN, topk=2),
         "small6": config1_params()}[cfg]
    ok = or_py.Keyframe.from_synth(kf)
    planes = random_planes(kf, 3, jitter=0.05)
    s32, v32 = or_py.score_field(ok, p, planes, views=True)
    s64, v64 = or_py.score_field_f64(ok, p, planes, views=True)
    m = np.isfinite(s32)
    dv = np.abs(v32[m] - v64[m])
    ds = np.abs(s32[m] - s64[m])
    print(f"{cfg}: per-view <=1e-5 {np.mean(dv <= 1e-5):.5f} max {dv.max():.2e}; "
          f"aggregated <=1e-5 {np.mean(ds <= 1e-5):.5f} max {ds.max():.2e}")
    assert np.mean(dv <= 1e-5) >= 0.999 and dv.max() <= 2e-3
    assert np.mean(ds <= 1e-5) >= 0.999


@pytest.mark.parametrize("cfg", ["c0", "small6"])
def test_score_error_split(cfg, kf_c0, kf_small6):
    """The FP32-vs-FP64 score error split into its two sources.
    (1) Warp: every tap's FP32 sample coordinates against the FP64 warp
        (exact bearings, the two-step transform of PatchMatchCuda.cu:117-130,
        libm atan2/asin) within 1e-3 px on the sphere: |dv|, and |du| cos(lat)
        (a longitude pixel spans cos(lat) of an equatorial one, so near the
        poles u alone is ill-conditioned while the image varies along u
        equally slowly).
    (2) NCC: the FP64 score with the samples taken at the FP32 coordinates
        (single-exp weights, raw moments) against the FP32 score: per view
        within 1e-5 on >= 99.998 % of (pixel, view) pairs (measured: 3 of
        273,504 small6 pairs exceed it, by 1.5e-5 at most, low-variance views
        whose FP32 moments cancel) and within 5e-5 everywhere; aggregated
        within 1e-5 on >= 99.999 % of pixels."""
    kf = kf_small6 if cfg == "small6" else kf_c0
    p = config1_params() if cfg == "small6" else config0_params()
    ok = or_py.Keyframe.from_synth(kf)
    planes = random_planes(kf, 3, jitter=0.05)
    h = p.patch_half
    ii, jj = np.mgrid[h + 1:kf.H - h - 1, h + 1:kf.W - h - 1]
    pix = np.stack([ii.ravel(), jj.ravel()], 1)
    uv32, uv64 = or_py.warp_taps(ok, p, planes, pix)
    du = (uv32[..., 0] - uv64[..., 0] + kf.W / 2) % kf.W - kf.W / 2
    dv = uv32[..., 1] - uv64[..., 1]
    coslat = np.cos(np.pi * (0.5 - uv64[..., 1] / kf.H))
    e = np.maximum(np.abs(du) * coslat, np.abs(dv))
    s32, v32 = or_py.score_field(ok, p, planes, views=True)
    s64, v64 = or_py.score_at_coords_f64(ok, p, planes, pix, uv32)
    dvw = np.abs(v32[pix[:, 0], pix[:, 1]] - v64)
    dag = np.abs(s32[pix[:, 0], pix[:, 1]] - s64)
    print(f"{cfg}: coordinates max {e.max():.2e} px (u alone {np.abs(du).max():.2e}); NCC given coordinates: "
          f"per view <=1e-5 {np.mean(dvw <= 1e-5):.7f} max {dvw.max():.2e}, aggregated <=1e-5 "
          f"{np.mean(dag <= 1e-5):.7f} max {dag.max():.2e}")
    assert e.max() <= 1e-3
    assert np.mean(dvw <= 1e-5) >= 0.99998 and dvw.max() <= 5e-5
    assert np.mean(dag <= 1e-5) >= 0.99999


def test_ground_truth_planes_score_high(kf_c0):
    ok = or_py.Keyframe.from_synth(kf_c0)
    s = or_py.score_field(ok, config0_params(agg_mode=AGG_MAX), gt_planes(kf_c0))
    assert np.nanmedian(s) > 0.9


# ------------------------------------------------------------------ init (PatchMatchCuda.cu:210-244)
def test_init_plane_identity_and_distribution(kf_c0):
    ok = or_py.Keyframe.from_synth(kf_c0)
    p = config0_params(coverage_mode=COVER_FULL)
    plane, depth, score = or_py.init(ok, p)
    B = synth.bearings(kf_c0.W, kf_c0.H)
    pb = np.sum(plane[..., :3].astype(np.float64) * B, -1)
    assert np.percentile(np.abs(pb * depth - 1), 99) < 1e-5          # p.(d b) = 1
    assert (pb > 0).mean() > 0.999                                   # normals face the camera
    inv = 1 / depth.astype(np.float64)
    assert inv.min() >= 1 / 20 - 1e-6 and inv.max() <= 1 / 0.3 + 1e-4
    h, _ = np.histogram(inv, bins=10, range=(1 / 20, 1 / 0.3))
    assert h.min() > 0.8 * h.mean()                                   # uniform in inverse depth
    pr = config0_params(init_mode=0, min_depth=1.0, max_depth=5.0)
    _, d2, _ = or_py.init(ok, pr)
    assert d2[:224, :].min() > 1.0 and d2.max() <= 5.0


def test_init_coverage_ref_grid(kf_c0):
    """480x240 with the reference launch grids (PatchMatchCuda.cu:538-539):
    rows >= 224 are never scored (score stays the NaN of the 0xFF memset)."""
    ok = or_py.Keyframe.from_synth(kf_c0)
    _, _, score = or_py.init(ok, config0_params())
    bits = score.view(np.uint32)
    assert (bits[224:, :] == 0xFFFFFFFF).all()
    assert np.isfinite(score[4:224, 4:476]).all()
    assert (bits[:3, :] == 0xFFFFFFFF).all() and (bits[:, :3] == 0xFFFFFFFF).all()


# ------------------------------------------------------------------ propagation invariants (PatchMatchCuda.cu:273-407)
def test_propagation_invariants(kf_c0):
    ok = or_py.Keyframe.from_synth(kf_c0)
    p = config0_params(iterations=2)
    plane, depth, score = or_py.init(ok, p)
    P1, D1, S1, ev = or_py.propagate(ok, p, plane, depth, score, 0, 1)
    # scores never decrease (strict-> acceptance)
    fin = np.isfinite(score)
    assert (S1[fin] >= score[fin]).all()
    # reference grid: rows >= 224 and columns >= 448 are never updated
    assert np.array_equal(P1[224:], plane[224:]) and np.array_equal(P1[:, 448:], plane[:, 448:])
    # output depth is 1/max(1e-6, |b.p|)
    B = synth.bearings(kf_c0.W, kf_c0.H)
    upd = np.any(P1 != plane, -1)
    bp = np.sum(P1[..., :3].astype(np.float64) * B, -1)
    dd = 1 / np.maximum(np.abs(bp), 1e-6)
    cond = np.linalg.norm(P1[..., :3].astype(np.float64), axis=-1) / np.maximum(np.abs(bp), 1e-30)
    assert (np.abs(dd[upd] / D1[upd] - 1) <= 1e-6 + 3e-7 * cond[upd]).all()
    assert upd.mean() > 0.5 and ev > 0
    # a pass touches only its colour: the (i+j)-even pass leaves every odd pixel alone
    Pr, Dr, Sr = plane.copy(), depth.copy(), score.copy()
    or_py.propagate_rows(ok, p, Pr, Dr, Sr, 0, 0, 0, kf_c0.H)
    ii, jj = np.indices(score.shape)
    odd = (ii + jj) % 2 == 1
    assert np.array_equal(Pr[odd], plane[odd]) and np.array_equal(Dr[odd], depth[odd])
    assert np.array_equal(Pr[~odd], P1[~odd])            # and equals the first half of the iteration


def test_finalize_window_and_postprocess(kf_c0):
    ok = or_py.Keyframe.from_synth(kf_c0)
    p = config0_params(iterations=1)
    out = or_py.densify(ok, p, kf_c0.mask)
    ng = out["nghbr"]
    # copy-back window [half, dim - half - 4) (depthmap.cc:731-732 with patch_size = half + 2)
    lo, hr, hc = 3, 240 - 7, 480 - 7
    win = np.zeros_like(ng, bool)
    win[lo:hr, lo:hc] = True
    assert np.array_equal(ng == 1, win & (kf_c0.mask > 0))
    # masked pixels: AssignMatrices values
    m0 = kf_c0.mask == 0
    assert (out["depth"][m0] == 0).all() and (out["score"][m0] == -1).all()
    # PostProcess: kept depths are the twice-median-filtered ones, rejected get (0, -1)
    kept = out["depth"] > 0
    assert (out["score"][~kept] == -1).all()
    assert kept.mean() > 0.3


def _postprocess(depth, score, **kw):
    """or_finalize with the same maps before and after the passes, no mask and
    no low-texture gate: pixels inside the copy-back window take the final
    maps, the others the initial ones, both (depth, score) here, so the
    output is PostProcess (depthmap.cc:1169-1208) of (depth, score) everywhere."""
    H, W = depth.shape
    img = np.zeros((H, W), np.uint8)
    pose = synth.yaw_pose(0.0, [0, 0, 0])
    kf = or_py.Keyframe(img, [img], pose, pose[None])
    p = default_params(patch_half=3, patch_stride=1, **kw)
    pl = np.zeros((H, W, 4), np.float32)
    maps = (pl, np.ascontiguousarray(depth, np.float32), np.ascontiguousarray(score, np.float32))
    od, _, osc, _ = or_py.finalize(kf, p, None, maps, maps)
    return od, osc


def test_postprocess_median_is_textbook_median_twice():
    """The kept depth is the 3x3 median with replicated border (cv::medianBlur,
    ksize 3) applied twice (depthmap.cc:1172-1173), against scipy's
    median_filter(size=3, mode='nearest') twice; the rejection clauses are
    switched off (threshold 1e300, depth range and score floor out of reach),
    so only d = 0 rejects.  One pass, or a zero border, would differ."""
    from scipy.ndimage import median_filter
    rng = np.random.default_rng(4)
    _, d = synth.render(synth.make_scene(8), synth.yaw_pose(0.3, [0.1, 0, 0]), 96, 48)
    d = d.astype(np.float32) * rng.choice([1.0, 1.0, 0.7, 1.4], d.shape).astype(np.float32)
    d[rng.random(d.shape) < 0.05] = 0.0
    sc = rng.uniform(-0.5, 1.0, d.shape).astype(np.float32)
    od, osc = _postprocess(d, sc, median_rel_thr=1e300, min_depth=-1e30, max_depth=1e30, min_score=-2.0)
    m2 = median_filter(median_filter(d, size=3, mode="nearest"), size=3, mode="nearest")
    keep = d != 0
    assert np.array_equal(od[keep].view(np.uint32), m2[keep].view(np.uint32))
    assert (od[~keep] == 0).all() and (osc[~keep] == -1).all()
    assert np.array_equal(osc[keep], sc[keep])
    m1 = median_filter(d, size=3, mode="nearest")
    assert np.mean(m1[keep] != m2[keep]) > 0.05
    mz = median_filter(median_filter(d, size=3, mode="constant"), size=3, mode="constant")
    border = np.zeros(d.shape, bool)
    border[[0, -1], :] = True
    border[:, [0, -1]] = True
    assert np.count_nonzero((mz != m2) & border & keep) > 10


# (m: the map's value around the pixel, d: the pixel's depth, sc: its score,
#  parameter overrides, kept?)  PostProcess (depthmap.cc:1189-1196) rejects when
#  d == 0, or fabs(d - m) / d > 0.05 with the float quotient against the double
#  literal, or d < min_depth, or d > max_depth, or sc <= min_score.
_F = np.float32
POSTPROCESS_TABLE = [
    ("plain", 19.0, 19.0, 0.5, {}, True),
    ("d = 0", 19.0, 0.0, 0.5, {}, False),
    # (20 - 19) / 20 rounds to 0.05f = 0.0500000007 > 0.05 (double): rejected
    ("ratio = 0.05f above", 19.0, 20.0, 0.5, {}, False),
    ("ratio = 0.05f below", 21.0, 20.0, 0.5, {}, False),
    # (20 - 19.000002) / 20 rounds below 0.05: kept (takes the median 19.000002)
    ("ratio just below 0.05", float(np.nextafter(_F(19.0), _F(20.0))), 20.0, 0.5, {}, True),
    ("ratio 0.06", 18.8, 20.0, 0.5, {}, False),
    ("d = min_depth", 19.0, 19.0, 0.5, {"min_depth": 19.0}, True),
    ("d < min_depth", 19.0, 19.0, 0.5, {"min_depth": float(np.nextafter(_F(19.0), _F(20.0)))}, False),
    ("d = max_depth", 19.0, 19.0, 0.5, {"max_depth": 19.0}, True),
    ("d > max_depth", 19.0, 19.0, 0.5, {"max_depth": float(np.nextafter(_F(19.0), _F(0.0)))}, False),
    ("score = min_score", 19.0, 19.0, float(_F(0.1)), {}, False),
    ("score just above min_score", 19.0, 19.0, float(np.nextafter(_F(0.1), _F(1.0))), {}, True),
    ("score -1", 19.0, 19.0, -1.0, {}, False),
    ("score NaN (never scored, 0xFF memset)", 19.0, 19.0, float("nan"), {}, True),
]


@pytest.mark.parametrize("case", POSTPROCESS_TABLE, ids=[c[0] for c in POSTPROCESS_TABLE])
def test_postprocess_truth_table(case):
    """One clause at a time: a pixel of depth d and score sc in a 32x32 map of
    value m (so both medians give m there); kept pixels take the median and
    keep their score, rejected ones become (0, -1)."""
    _, m, d, sc, kw, kept = case
    depth = np.full((32, 32), m, np.float32)
    score = np.full((32, 32), 0.5, np.float32)
    depth[16, 16] = d
    score[16, 16] = sc
    kw = dict(dict(min_depth=1e-5, max_depth=100.0, min_score=0.1), **kw)
    od, osc = _postprocess(depth, score, **kw)
    if kept:
        assert od[16, 16] == _F(m) and (osc[16, 16] == _F(sc) or (np.isnan(sc) and np.isnan(osc[16, 16])))
    else:
        assert od[16, 16] == 0.0 and osc[16, 16] == -1.0


# ------------------------------------------------------------------ committed fixture
def test_golden_fixture_reproduced():
    from tests_golden_loader import load_fixture
    fx, p = load_fixture()
    kf = or_py.Keyframe(fx["ref"], [fx["view0"], fx["view1"]], fx["ref_pose"], fx["view_poses"])
    out = or_py.densify(kf, p, fx["mask"])
    for k in ("depth", "plane", "score", "nghbr", "raw_plane", "raw_depth", "raw_score"):
        assert np.array_equal(np.asarray(out[k]).view(np.uint8), np.asarray(fx[k]).view(np.uint8)), k


# ------------------------------------------------------------------ coefficient tables
def test_polynomial_coefficients_match_fit_and_each_other():
    """The oracle's and the device's tables are the same float32 literals, and
    they are the output of patchmatch_amd/csrc/gen_coeffs.py (within 2 ulp: the
    least-squares solve may differ in the last bit across BLAS builds)."""
    import importlib.util
    import re
    root = os.path.dirname(HERE)
    spec = importlib.util.spec_from_file_location("gen_coeffs", os.path.join(root, "patchmatch_amd", "csrc",
                                                                             "gen_coeffs.py"))
    gc = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gc)
    fit = gc.coefficients()
    lit = re.compile(r"-?\d\.\d{9}e[+-]\d{2}f")
    dev = [float(x[:-1]) for x in lit.findall(open(os.path.join(root, "patchmatch_amd", "csrc", "pmp_math.h")).read())]
    orc = [float(x[:-1]) for x in lit.findall(open(os.path.join(root, "oracle", "or_math.h")).read())]
    assert dev == orc and len(dev) >= 30
    table = np.float32(dev)
    for name, v in fit.items():
        for x in v:
            if abs(x) in (0.0, 1.0, 0.5):
                continue
            near = np.min(np.abs(table - x) / np.spacing(np.float32(abs(x))))
            assert near <= 2, (name, x)
