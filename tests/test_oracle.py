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
from patchmatch_amd.params import (AGG_MAX, AGG_REF_FOLD, AGG_TOPK_MEAN, COVER_FULL, EDGE_WRAP, config0_params,
                                   config1_params, default_params)

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
    scale exactly with the binade; monotone rounding covers mn < mx), and
    m rsqrt(m M) exceeds 1 only for M = m, by one ulp."""
    m = (np.arange(1 << 23, dtype=np.uint32) | np.uint32(0x3F800000)).view(np.float32)
    assert (m * or_py.eval_math("rcp", m)).astype(np.float32).max() <= 1.0
    for lo in (0x3F800000, 0x40000000):     # rsqrt repeats every two binades
        m = (np.arange(1 << 23, dtype=np.uint32) | np.uint32(lo)).view(np.float32)
        nxt = np.nextafter(m, np.float32(np.inf))
        assert (m * or_py.eval_math("rsqrt", (m * nxt).astype(np.float32))).astype(np.float32).max() <= 1.0
        assert (m * or_py.eval_math("rsqrt", (m * m).astype(np.float32))).astype(np.float32).max() \
            <= np.float32(1.0) + np.float32(2.0 ** -23)
    # both poles and the horizon without a clamp on h^2: (b, a, b) with b = 0
    lat = or_py.eval_math("lat", np.float32([2.0, -3.0, 0.0]), np.float32([0.0, 0.0, 0.0]))
    assert lat[0] == np.float32(np.pi / 2) and lat[1] == -np.float32(np.pi / 2) and lat[2] == 0.0


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


def test_view_aggregation_truth_tables():
    fold = default_params(agg_mode=AGG_REF_FOLD)
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


# ------------------------------------------------------------------ FP32 vs FP64 (fixed hypothesis field)
@pytest.mark.parametrize("cfg", ["c0", "small6"])
def test_fp32_score_vs_fp64_oracle(cfg, kf_c0, kf_small6):
    """The NCC of a fixed hypothesis field against the FP64 scorer (exact
    bearings, two-step transform, libm functions, single-exp weight, raw
    moments).  Per view: |d| <= 1e-5 on >= 99 % of interior pixels and
    <= 2e-3 everywhere; aggregated (best-k mean): <= 1e-5 on >= 99 %."""
    kf = kf_c0 if cfg == "c0" else kf_small6
    p = config0_params(agg_mode=AGG_TOPK_MEAN, topk=2) if cfg == "c0" else config1_params()
    ok = or_py.Keyframe.from_synth(kf)
    planes = random_planes(kf, 3, jitter=0.05)
    s32, v32 = or_py.score_field(ok, p, planes, views=True)
    s64, v64 = or_py.score_field_f64(ok, p, planes, views=True)
    m = np.isfinite(s32)
    dv = np.abs(v32[m] - v64[m])
    ds = np.abs(s32[m] - s64[m])
    print(f"{cfg}: per-view <=1e-5 {np.mean(dv <= 1e-5):.5f} max {dv.max():.2e}; "
          f"aggregated <=1e-5 {np.mean(ds <= 1e-5):.5f} max {ds.max():.2e}")
    assert np.mean(dv <= 1e-5) >= 0.99 and dv.max() <= 2e-3
    assert np.mean(ds <= 1e-5) >= 0.99


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
