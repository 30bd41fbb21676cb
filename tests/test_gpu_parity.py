"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

Integer/index work and the canonical FP32 arithmetic are compared bit for
bit (the oracle executes the same IEEE operation sequence); the parity bar
of the deterministic mode, >= 99.9 % of valid pixels within 1e-4 relative
depth, is asserted on top of exact equality statistics that are printed.
"""
import numpy as np
import pytest

from oracle import or_py
from patchmatch_amd._lib import PMP_MAPS_FINAL
from patchmatch_amd.params import (AGG_FOLD_INDEP, AGG_MAX, AGG_REF_FOLD, AGG_TOPK_MEAN, COVER_FULL, EDGE_CLAMP,
                                   EDGE_ZERO, config0_params, config1_params, default_params)

from conftest import cached, random_planes

pytestmark = pytest.mark.gpu


def _same_bits(a, b):
    a = np.ascontiguousarray(a, np.float32).view(np.uint32)
    b = np.ascontiguousarray(b, np.float32).view(np.uint32)
    return a == b


def _report(name, a, b):
    eq = _same_bits(a, b)
    print(f"[{name}] bit-equal {eq.mean() * 100:.4f}% of {eq.size}", flush=True)
    return eq


MATH_CASES = {
    "rcp": lambda r: np.concatenate([r.uniform(-1e4, 1e4, 200000), r.uniform(-1, 1, 100000),
                                     [0.0, -0.0, 1e-31, -1e-31, 1e31, np.inf, -np.inf, np.nan, 1e-6]]),
    "rsqrt": lambda r: np.concatenate([10 ** r.uniform(-12, 12, 300000), [1e-31, 1e31, 1.0, 4.0]]),
    "sqrt": lambda r: np.concatenate([10 ** r.uniform(-12, 12, 300000), [0.0, 1e-31, 1e31, -1.0, np.nan]]),
    "asin": lambda r: np.concatenate([r.uniform(-1, 1, 300000), [-1.0, 1.0, 0.5, -0.5, 0.0, 2.0, np.nan]]),
    "acos": lambda r: np.concatenate([r.uniform(-1, 1, 300000), [-1.0, 1.0, 0.5, -0.5, 0.0]]),
    "exp": lambda r: np.concatenate([r.uniform(-90, 90, 300000), r.uniform(-0.2, 0.2, 100000),
                                     [0.0, -87.3, 88.0, -np.inf, np.nan]]),
    "log": lambda r: ((r.integers(0, 1 << 24, 300000) + 1) * 2.0 ** -24),
    "sin2pi": lambda r: (r.integers(0, 1 << 24, 300000) * 2.0 ** -24),
    "cos2pi": lambda r: (r.integers(0, 1 << 24, 300000) * 2.0 ** -24),
}


@pytest.mark.parametrize("name", sorted(MATH_CASES) + ["atan2", "lat", "lon"])
def test_math_bitwise(gpu_ctx, name):
    rng = np.random.default_rng(11)
    if name == "atan2":
        a = np.concatenate([rng.normal(0, 10, 300000), [0.0, 0.0, -1.0, 1e-35, 5.0, -0.0, -0.0],
                            [1e30, -1e-40, 3e29, -1e30]]).astype(np.float32)
        b = np.concatenate([rng.normal(0, 10, 300000), [0.0, -1.0, 0.0, 1e-35, -5.0, -1.0, 1.0],
                            [1e-40, 1e30, -3e29, -1e30]]).astype(np.float32)
    elif name == "lat":    # latitude of the direction (b, a, b)
        a = np.concatenate([rng.normal(0, 10, 300000), rng.normal(0, 1e-3, 1000),
                            [0.0, -0.0, 1.0, -1.0, 1e-20, 7.0]]).astype(np.float32)
        b = np.concatenate([rng.normal(0, 10, 300000), rng.normal(0, 10, 1000),
                            [0.0, 1.0, 0.0, 0.0, 1e-20, 7.0 / np.sqrt(2)]]).astype(np.float32)
    elif name == "lon":    # longitude of the direction (a, 0, b)
        a = np.concatenate([rng.normal(0, 10, 300000), rng.uniform(-1, 1, 1000),
                            [0.0, 0.0, -0.0, 1.0, -1.0, 1e-20, 3.0, -3.0]]).astype(np.float32)
        b = np.concatenate([rng.normal(0, 10, 300000), rng.uniform(-1e-3, 1e-3, 1000),
                            [0.0, -1.0, -1.0, 0.0, 0.0, 1e-20, 3.0, -3.0]]).astype(np.float32)
    else:
        a = MATH_CASES[name](rng).astype(np.float32)
        b = None
    g = gpu_ctx.eval_math(name, a, b)
    o = or_py.eval_math(name, a, b)
    eq = _same_bits(g, o) | (np.isnan(g) & np.isnan(o))
    assert eq.all(), f"{name}: {np.count_nonzero(~eq)} mismatches, e.g. x={a[~eq][:4]} gpu={g[~eq][:4]} cpu={o[~eq][:4]}"


def test_refine_normals_bitwise(gpu_ctx, kf_c0):
    ok = or_py.Keyframe.from_synth(kf_c0)
    p = config0_params(seed=0x1234567890ABCDEF)
    for (i, j, it) in [(0, 0, 0), (10, 20, 3), (239, 479, 7), (100, 101, 1)]:
        g = gpu_ctx.refine_normals(p, kf_c0.W, i, j, it, 18)
        o = or_py.refine_normals(ok, p, i, j, it, 18)
        assert _same_bits(g, o).all(), (i, j, it, g, o)


def _load(ctx, kf):
    ctx.load_keyframe(kf.ref, kf.views, kf.ref_pose, kf.view_poses)


@pytest.mark.parametrize("pname", ["c0", "c0_fold_ref_init", "c0_max_clamp", "c0_zero_edge"])
def test_init_bitwise(gpu_ctx, kf_c0, pname):
    p = {"c0": config0_params(),
         "c0_fold_ref_init": config0_params(init_mode=0, agg_mode=AGG_REF_FOLD, max_depth=8.0, seed=7),
         "c0_max_clamp": config0_params(agg_mode=AGG_MAX, edge_mode=EDGE_CLAMP, seed=3),
         "c0_zero_edge": config0_params(edge_mode=EDGE_ZERO, coverage_mode=COVER_FULL, seed=5)}[pname]
    _load(gpu_ctx, kf_c0)
    print(f"case {pname}", flush=True)
    gpu_ctx.random_initialization(p)
    gp, gd, gs = gpu_ctx.get_maps(0)
    op, od, os_ = or_py.init(or_py.Keyframe.from_synth(kf_c0), p)
    assert _report("plane", gp, op).all()
    assert _report("depth", gd, od).all()
    assert _report("score", gs, os_).all()
    ip, idp, isc = gpu_ctx.get_maps(1)
    assert _same_bits(ip, gp).all() and _same_bits(idp, gd).all() and _same_bits(isc, gs).all()


@pytest.mark.parametrize("agg", [AGG_REF_FOLD, AGG_MAX, AGG_TOPK_MEAN, AGG_FOLD_INDEP])
def test_score_field_bitwise(gpu_ctx, kf_small6, agg):
    kf = kf_small6
    p = config1_params(agg_mode=agg, coverage_mode=COVER_FULL)
    _load(gpu_ctx, kf)
    planes = random_planes(kf, 5)
    gpu_ctx.set_maps(planes, np.ones((kf.H, kf.W), np.float32), np.zeros((kf.H, kf.W), np.float32))
    gpu_ctx.score_plane_depth(p)
    _, _, gs = gpu_ctx.get_maps(0)
    os_ = or_py.score_field(or_py.Keyframe.from_synth(kf), p, planes)
    ok = ~np.isnan(os_)
    assert _report("score", gs[ok], os_[ok]).all()


@pytest.mark.parametrize("edge", [EDGE_CLAMP, EDGE_ZERO])
def test_init_and_score_ragged_tiles(gpu_ctx, edge):
    """Init and score kernels stage 4 x 64 tiles of the reference with a halo
    of the patch half-width: at 136 x 70 (a partial tile column of 8 pixels,
    a partial tile row of 2) with the largest half-width (11 x 11 at stride
    2), every map bit for bit against the oracle."""
    from patchmatch_amd import synth
    kf = cached("ragged", lambda: synth.config1(4, W=136, H=70, n_views=3, base_freq=synth.TEX_FREQ_C0))
    p = config1_params(coverage_mode=COVER_FULL, edge_mode=edge, seed=11)
    assert p.patch_half * p.patch_stride == 10
    _load(gpu_ctx, kf)
    print(f"case ragged edge {edge}", flush=True)
    gpu_ctx.random_initialization(p)
    gp, gd, gs = gpu_ctx.get_maps(0)
    op, od, os_ = or_py.init(or_py.Keyframe.from_synth(kf), p)
    assert _report("plane", gp, op).all()
    assert _report("depth", gd, od).all()
    assert _report("score", gs, os_).all()
    planes = random_planes(kf, 9)
    gpu_ctx.set_maps(planes, np.ones((kf.H, kf.W), np.float32), np.zeros((kf.H, kf.W), np.float32))
    gpu_ctx.score_plane_depth(p)
    _, _, gs = gpu_ctx.get_maps(0)
    os_ = or_py.score_field(or_py.Keyframe.from_synth(kf), p, planes)
    ok = ~np.isnan(os_)
    assert _report("score field", gs[ok], os_[ok]).all()


def test_propagate_c0(gpu_ctx, kf_c0):
    """Config 0 (480x240, 2 views, 7x7 patch, the reference's pooled fold,
    reference launch grids): init + 4 red-black iterations, every map bit for
    bit after every iteration (the north star's >= 99.9 % within 1e-4 relative
    depth follows)."""
    p = config0_params()
    assert p.agg_mode == AGG_REF_FOLD and p.coverage_mode == 0
    _load(gpu_ctx, kf_c0)
    gpu_ctx.random_initialization(p)
    ok = or_py.Keyframe.from_synth(kf_c0)
    op, od, os_ = or_py.init(ok, p)
    for it in range(p.iterations):
        gpu_ctx.red_black_pass(p, it, it + 1)
        op, od, os_, _ = or_py.propagate(ok, p, op, od, os_, it, it + 1)
        gp, gd, gs = gpu_ctx.get_maps(0)
        for name, g, o in (("depth", gd, od), ("plane", gp, op), ("score", gs, os_)):
            eq = _report(f"it{it} {name}", g, o) | (np.isnan(g) & np.isnan(o))
            assert eq.all(), f"iteration {it} {name}: {np.count_nonzero(~eq)} of {eq.size} differ"
    # checkerboard/coverage/mask: the unscored pixels (NaN scores) are the same set
    assert np.array_equal(np.isnan(gs), np.isnan(os_))


@pytest.mark.parametrize("edge", [EDGE_CLAMP, EDGE_ZERO])
def test_propagate_largest_patch(gpu_ctx, edge):
    """The largest patch (half 10, stride 2: 121 taps, the widest halo of the
    per-tile LDS copies) on a 200x100 keyframe whose width is not a multiple
    of the 128-column tiles: init and one red-black iteration bit for bit."""
    from patchmatch_amd import synth
    kf = synth.config1(9, W=200, H=100, n_views=3)
    p = config1_params(patch_half=10, patch_stride=2, iterations=1, edge_mode=edge, seed=31)
    _load(gpu_ctx, kf)
    gpu_ctx.random_initialization(p)
    gpu_ctx.red_black_pass(p, 0, 1)
    ok = or_py.Keyframe.from_synth(kf)
    op, od, os_ = or_py.init(ok, p)
    d_init = od.copy()
    op, od, os_, _ = or_py.propagate(ok, p, op, od, os_, 0, 1)
    gp, gd, gs = gpu_ctx.get_maps(0)
    for name, g, o in (("plane", gp, op), ("depth", gd, od), ("score", gs, os_)):
        eq = _report(f"half 10 edge {edge} {name}", g, o) | (np.isnan(g) & np.isnan(o))
        assert eq.all(), f"{name}: {np.count_nonzero(~eq)} of {eq.size} differ"
    assert np.count_nonzero(od != d_init) > 1000   # the pass updated pixels


@pytest.mark.parametrize("n_views,agg", [(1, AGG_TOPK_MEAN), (4, AGG_REF_FOLD), (10, AGG_TOPK_MEAN), (10, AGG_MAX),
                                         (10, AGG_REF_FOLD), (5, AGG_FOLD_INDEP)])
def test_propagate_view_counts(gpu_ctx, n_views, agg):
    """View counts that split into groups of 3, 2 and 1 (1; 3+1; 3+2;
    3+3+3+1), the maximum of 10 views (the pooled fold carries its running
    sums across groups), and a top-3 mean over a single view (k clamped to N),
    on a 150x74 keyframe (neither dimension a multiple of a tile): init and one
    red-black iteration bit for bit."""
    from patchmatch_amd import synth
    kf = synth.make_keyframe(20 + n_views, 150, 74, n_views, baseline=(0.05, 0.5), spheres=6, tube=True)
    p = config1_params(agg_mode=agg, iterations=1, seed=7)
    _load(gpu_ctx, kf)
    gpu_ctx.random_initialization(p)
    gpu_ctx.red_black_pass(p, 0, 1)
    ok = or_py.Keyframe.from_synth(kf)
    op, od, os_ = or_py.init(ok, p)
    d_init = od.copy()
    op, od, os_, _ = or_py.propagate(ok, p, op, od, os_, 0, 1)
    gp, gd, gs = gpu_ctx.get_maps(0)
    for name, g, o in (("plane", gp, op), ("depth", gd, od), ("score", gs, os_)):
        eq = _report(f"{n_views} views agg {agg} {name}", g, o) | (np.isnan(g) & np.isnan(o))
        assert eq.all(), f"{name}: {np.count_nonzero(~eq)} of {eq.size} differ"
    assert np.count_nonzero(od != d_init) > 500


@pytest.mark.parametrize("it", [0, 5])
def test_propagate_full_size_row_band(gpu_ctx, it):
    """The 1920x960 keyframe with 6 neighbours at full size: the GPU's
    iteration `it` (the first, and the sixth of eight) from its own maps before
    it, against the oracle recomputing a band of rows from the same maps (red
    pass on the band plus the 3-row reach of the candidates, then the black
    pass on the band).  Rows of one colour only read the other colour within
    a pass, so the band equals the whole pass there: bit for bit."""
    from patchmatch_amd import synth
    kf = cached("c1_full", lambda: synth.config1(0, W=1920, H=960, n_views=6))
    p = config1_params(seed=0)
    _load(gpu_ctx, kf)
    gpu_ctx.random_initialization(p)
    if it:
        gpu_ctx.red_black_pass(p, 0, it)
    ip, idp, isc = gpu_ctx.get_maps(0)
    gpu_ctx.red_black_pass(p, it, it + 1)
    gp, gd, gs = gpu_ctx.get_maps(0)
    ok = or_py.Keyframe.from_synth(kf)
    op, od, os_ = (np.array(a, copy=True) for a in (ip, idp, isc))
    r0, r1 = 470, 506
    or_py.propagate_rows(ok, p, op, od, os_, it, 0, r0 - 3, r1 + 3)
    or_py.propagate_rows(ok, p, op, od, os_, it, 1, r0, r1)
    band = slice(r0, r1)
    for name, g, o in (("plane", gp[band], op[band]), ("depth", gd[band], od[band]), ("score", gs[band], os_[band])):
        eq = _report(f"1920x960 iteration {it} rows {r0}-{r1} {name}", g, o) | (np.isnan(g) & np.isnan(o))
        assert eq.all(), f"{name}: {np.count_nonzero(~eq)} of {eq.size} differ"
    assert np.count_nonzero(od[band] != idp[band]) > (0.3 if it == 0 else 0.02) * od[band].size


def test_propagate_largest_served_size_row_band(gpu_ctx):
    """The largest keyframes of the served stream, 5760x2880 with 4 neighbours
    from one trajectory (rendered on the GPU): the first iteration on a
    12-row band recomputed by the oracle, bit for bit.  Covers the 24-bit
    index products and byte offsets at the largest width and height."""
    from patchmatch_amd import synth
    seq = synth.sequence_gpu(17, 5, 5760, 2880, with_gt=False)
    ims, poses = seq["images"], seq["poses"]
    kf = synth.Keyframe(W=5760, H=2880, ref=ims[2], views=[ims[k] for k in (0, 1, 3, 4)],
                        ref_pose=poses[2], view_poses=np.stack([poses[k] for k in (0, 1, 3, 4)]),
                        gt_depth=None, mask=seq["mask"], bgr=None)
    p = config1_params(seed=5)
    _load(gpu_ctx, kf)
    gpu_ctx.random_initialization(p)
    ip, idp, isc = gpu_ctx.get_maps(0)
    gpu_ctx.red_black_pass(p, 0, 1)
    gp, gd, gs = gpu_ctx.get_maps(0)
    ok = or_py.Keyframe.from_synth(kf)
    op, od, os_ = (np.array(a, copy=True) for a in (ip, idp, isc))
    r0, r1 = 2000, 2012
    or_py.propagate_rows(ok, p, op, od, os_, 0, 0, r0 - 3, r1 + 3)
    or_py.propagate_rows(ok, p, op, od, os_, 0, 1, r0, r1)
    band = slice(r0, r1)
    for name, g, o in (("plane", gp[band], op[band]), ("depth", gd[band], od[band]), ("score", gs[band], os_[band])):
        eq = _report(f"5760x2880 rows {r0}-{r1} {name}", g, o) | (np.isnan(g) & np.isnan(o))
        assert eq.all(), f"{name}: {np.count_nonzero(~eq)} of {eq.size} differ"
    assert np.count_nonzero(od[band] != idp[band]) > 0.3 * od[band].size


def test_plain_rejects_bad_params_and_runs_smallest(gpu_ctx):
    """Argument errors of the one-keyframe path (patch half, view count,
    iterations, image below the patch) come back as EINVAL, and the smallest
    image the 7x7 patch allows (14x14) densifies bit for bit as the oracle."""
    from patchmatch_amd import synth
    from patchmatch_amd._lib import PmpError
    kf = synth.make_keyframe(31, 32, 16, 2)
    _load(gpu_ctx, kf)
    for bad in (config0_params(patch_half=11), config0_params(iterations=65), config0_params(patch_half=2)):
        with pytest.raises(PmpError, match="EINVAL"):
            gpu_ctx.densify(bad)
    with pytest.raises(PmpError, match="EINVAL"):
        gpu_ctx.set_views(0, kf.ref_pose, [1] * 11, [kf.view_poses[0]] * 11)
    with pytest.raises(PmpError, match="EINVAL"):   # 16 rows < 2 * 5 + 8
        gpu_ctx.densify(config0_params(patch_half=5))
    small = synth.make_keyframe(32, 14, 14, 2)
    p = config0_params(iterations=2, coverage_mode=COVER_FULL)
    _load(gpu_ctx, small)
    gpu_ctx.densify(p, small.mask)
    _, gd, gs, gn = gpu_ctx.get_maps(PMP_MAPS_FINAL)
    ref = or_py.densify(or_py.Keyframe.from_synth(small), p, small.mask)
    assert _same_bits(gd, ref["depth"]).all() and _same_bits(gs, ref["score"]).all()
    assert np.array_equal(gn, ref["nghbr"])


def test_densify_c0_final(gpu_ctx, kf_c0):
    p = config0_params()
    _load(gpu_ctx, kf_c0)
    gpu_ctx.densify(p, kf_c0.mask)
    gp, gd, gs, gn = gpu_ctx.get_maps(2)
    out = or_py.densify(or_py.Keyframe.from_synth(kf_c0), p, kf_c0.mask)
    assert np.array_equal(gn, out["nghbr"])
    assert np.array_equal(gd == 0, out["depth"] == 0)
    assert _report("final depth", gd, out["depth"]).all()
    assert _report("final score", gs, out["score"]).all()
    assert _report("final plane", gp[..., :3], out["plane"]).all()


@pytest.mark.parametrize("sd,post", [(0.0, 1), (6.0, 1), (6.0, 0)])
def test_finalize_crafted_bitwise(gpu_ctx, kf_c0, sd, post):
    """pmp_finalize (copy-back window, low-texture gate, PostProcess) on crafted
    final maps holding the PostProcess truth-table values (ratios exactly at
    0.05f on both sides and just below, zero depths, depths at and beyond the
    range ends, scores at and just above min_score, -1 and NaN) over a noisy
    depth field, against or_finalize bit for bit; with min_patch_sd > 0 the
    gate (variance of the tap samples below sd^2, depthmap.cc:739-743) zeroes
    low-texture pixels, which the run without PostProcess shows directly."""
    p = config0_params(min_patch_sd=sd, postprocess=post, min_depth=1e-5, max_depth=30.0)
    _load(gpu_ctx, kf_c0)
    gpu_ctx.random_initialization(p)
    ip, idp, isc = gpu_ctx.get_maps(1)
    H, W = kf_c0.H, kf_c0.W
    rng = np.random.default_rng(17)
    depth = (19.0 * rng.choice([1.0, 1.0, 1.0, 1.03, 0.9], (H, W))).astype(np.float32)
    score = rng.uniform(-0.2, 1.0, (H, W)).astype(np.float32)
    f = np.float32
    specials = [(19.0, 20.0, 0.5), (21.0, 20.0, 0.5), (float(np.nextafter(f(19.0), f(20.0))), 20.0, 0.5),
                (19.0, 0.0, 0.5), (30.0, 30.0, 0.5), (30.0, float(np.nextafter(f(30.0), f(31.0))), 0.5),
                (19.0, 19.0, float(f(0.1))), (19.0, 19.0, float(np.nextafter(f(0.1), f(1.0)))),
                (19.0, 19.0, -1.0), (19.0, 19.0, float("nan"))]
    for k, (m, d, sc) in enumerate(specials):
        i, j = 40 + 6 * (k % 5), 60 + 40 * (k // 5)
        depth[i - 2:i + 3, j - 2:j + 3] = m
        depth[i, j] = d
        score[i, j] = sc
    plane = np.zeros((H, W, 4), np.float32)
    plane[..., :3] = rng.normal(size=(H, W, 3)).astype(np.float32)
    gpu_ctx.set_maps(plane, depth, score)
    gpu_ctx.finalize(p, kf_c0.mask)
    gp, gd, gs, gn = gpu_ctx.get_maps(PMP_MAPS_FINAL)
    ok = or_py.Keyframe.from_synth(kf_c0)
    od, op, osc, on = or_py.finalize(ok, p, kf_c0.mask, (ip, idp, isc), (plane, depth, score))
    print(f"min_patch_sd {sd} postprocess {post}: kept {np.count_nonzero(od)} of {od.size}", flush=True)
    assert _report("depth", gd, od).all()
    eq = _report("score", gs, osc) | (np.isnan(gs) & np.isnan(osc))
    assert eq.all()
    assert _same_bits(gp[..., :3], op).all() and np.array_equal(gn, on)
    if sd > 0 and not post:
        gated = (on == 1) & (osc == 0.0) & (od == 0.0)
        print(f"gated (low texture) {np.count_nonzero(gated)}", flush=True)
        assert np.count_nonzero(gated) > 100


def test_normals_export(gpu_ctx, kf_c0):
    """pmp_get_normals: -p/|p| of the final planes within 2 ulp of float64,
    zero where the plane is zero; n . b has the opposite sign of p . b, so the
    normal faces the camera wherever the plane lies in front of it."""
    from patchmatch_amd import synth
    p = config0_params(iterations=1)
    _load(gpu_ctx, kf_c0)
    gpu_ctx.densify(p, kf_c0.mask)
    plane = gpu_ctx.get_maps(PMP_MAPS_FINAL)[0][..., :3].astype(np.float64)
    n = gpu_ctx.get_normals(PMP_MAPS_FINAL).astype(np.float64)
    norm = np.linalg.norm(plane, axis=-1)
    has = norm > 0
    want = -plane[has] / norm[has][:, None]
    assert np.max(np.abs(n[has] - want)) <= 2.5e-7
    assert np.all(n[~has] == 0.0)
    B = synth.bearings(kf_c0.W, kf_c0.H)
    pb = np.sum(plane[has] * B[has], axis=-1)
    nb = np.sum(n[has] * B[has], axis=-1)
    clear = np.abs(pb) > 1e-6 * norm[has]
    assert np.all(np.sign(nb[clear]) == -np.sign(pb[clear]))
    assert np.mean(pb > 0) > 0.99 and has.mean() > 0.5


def test_clean_bitwise(gpu_ctx, kf_c0):
    from patchmatch_amd import synth
    rng = np.random.default_rng(2)
    traj = synth.trajectory(1, 5, 0.15)
    scene = synth.make_scene(42)
    depths = []
    for pose in traj:
        _, d = synth.render(scene, pose, 240, 120)
        d = d.astype(np.float32) * rng.choice([1.0, 1.0, 1.0, 1.2], d.shape).astype(np.float32)
        d[rng.random(d.shape) < 0.05] = 0.0
        depths.append(d)
    g = gpu_ctx.clean(depths, traj, 0.03, 3)
    o = or_py.clean(depths, traj, 0.03, 3)
    print(f"kept {np.count_nonzero(o)} of {o.size}")
    assert _same_bits(g, o).all()


@pytest.mark.parametrize("px,deg", [(0.5, 0.0), (0.0, 10.0), (0.75, 12.5)])
def test_clean_checks_bitwise(gpu_ctx, px, deg):
    """pmp_clean_ex (round trip within px pixels, normals within deg degrees)
    against the oracle on noisy maps with a third of the planes tilted 25 degrees."""
    from patchmatch_amd import synth
    rng = np.random.default_rng(3)
    traj = synth.trajectory(4, 5, 0.15)
    scene = synth.make_scene(45, spheres=3)
    depths, planes = [], []
    for pose in traj:
        _, d = synth.render(scene, pose, 240, 120)
        d = d.astype(np.float32) * rng.choice([1.0, 1.0, 1.0, 1.02], d.shape).astype(np.float32)
        d[rng.random(d.shape) < 0.05] = 0.0
        depths.append(d)
        planes.append(synth.tilt_planes(synth.render_planes(scene, pose, 240, 120), rng, 0.33, np.radians(25.0)))
    g = gpu_ctx.clean_ex(depths, planes, traj, 0.03, 3, px, deg)
    o = or_py.clean_ex(depths, planes, traj, 0.03, 3, px, deg)
    plain = or_py.clean(depths, traj, 0.03, 3)
    print(f"px {px} deg {deg}: kept {np.count_nonzero(o)} of {np.count_nonzero(plain)} without the tests", flush=True)
    assert np.count_nonzero(o) < np.count_nonzero(plain)
    assert _same_bits(g, o).all()
    off = gpu_ctx.clean_ex(depths, None, traj, 0.03, 3, 0.0, 0.0)
    assert _same_bits(off, plain).all()


def test_clean_checks_reject_bad_arguments(gpu_ctx):
    from patchmatch_amd._lib import PmpError
    d = np.ones((40, 80), np.float32)
    pose = np.concatenate([np.eye(3).ravel(), np.zeros(3)])
    with pytest.raises(PmpError, match="EINVAL"):
        gpu_ctx.clean_ex([d, d], None, [pose, pose], 0.01, 2, 0.0, 10.0)       # normals without planes
    pl = np.zeros((40, 80, 4), np.float32)
    with pytest.raises(PmpError, match="EINVAL"):
        gpu_ctx.clean_ex([d, d], [pl, pl], [pose, pose], 0.01, 2, 0.0, 90.0)    # no angle bound at 90 degrees
    with pytest.raises(PmpError, match="EINVAL"):
        gpu_ctx.clean_ex([d, d], None, [pose, pose], 0.01, 2, float("nan"), 0.0)


def test_prune_sequence_bitwise(gpu_ctx):
    """Five queued maps, three prunes with pops: pruned maps, modified later maps
    and emitted points equal the oracle's raster-order loop exactly."""
    from patchmatch_amd import synth
    rng = np.random.default_rng(9)
    traj = synth.trajectory(2, 8, 0.1)
    scene = synth.make_scene(43)
    maps, bgrs = [], []
    for pose in traj:
        img, d = synth.render(scene, pose, 200, 100)
        d = d.astype(np.float32) * rng.choice([1.0, 1.0, 1.005, 0.97], d.shape).astype(np.float32)
        d[rng.random(d.shape) < 0.03] = 0.0
        maps.append(d)
        bgrs.append(np.repeat(np.clip(img, 0, 255).astype(np.uint8)[..., None], 3, 2))
    gpu_ctx.pruner_reset()
    cpu_maps = [m.copy() for m in maps]
    for k in range(5):
        gpu_ctx.pruner_add_view(maps[k], traj[k], bgrs[k])
    queue = list(range(5))
    nxt = 5
    for step in range(3):
        gpr, gpts, gcols, gkps = gpu_ctx.pruner_prune(0.03, maps[0].shape)
        opr, opts, ocols, okps = or_py.prune([cpu_maps[q] for q in queue], traj[queue], bgrs[queue[0]], 0.03)
        print(f"step {step}: {len(opts)} points", flush=True)
        assert _same_bits(gpr, opr).all()
        assert np.array_equal(gkps, okps) and np.array_equal(gcols, ocols)
        assert np.array_equal(gpts.view(np.uint64), opts.view(np.uint64))
        for idx, q in enumerate(queue):
            assert _same_bits(gpu_ctx.pruner_get_map(idx, maps[0].shape), cpu_maps[q]).all()
        gpu_ctx.pruner_pop_head()
        queue.pop(0)
        gpu_ctx.pruner_add_view(maps[nxt], traj[nxt], bgrs[nxt])
        queue.append(nxt)
        nxt += 1


def test_golden_fixture_on_gpu(gpu_ctx):
    from tests_golden_loader import load_fixture
    fx, p = load_fixture()
    gpu_ctx.load_keyframe(fx["ref"], [fx["view0"], fx["view1"]], fx["ref_pose"], fx["view_poses"])
    gpu_ctx.densify(p, fx["mask"])
    gp, gd, gs, gn = gpu_ctx.get_maps(2)
    assert _same_bits(gd, fx["depth"]).all() and _same_bits(gs, fx["score"]).all()
    assert _same_bits(gp[..., :3], fx["plane"]).all() and np.array_equal(gn, fx["nghbr"])
    wp, wd, ws = gpu_ctx.get_maps(0)
    assert _same_bits(wd, fx["raw_depth"]).all() and _same_bits(wp, fx["raw_plane"]).all()


def test_host_mask_changes_between_calls(gpu_ctx, kf_c0):
    """A host mask is uploaded through a pinned copy and skipped when its bytes
    equal the last upload's: finalize with mask A, then B (same size, other
    bytes), then A again must give what the pooled masks A and B give."""
    p = config0_params()
    _load(gpu_ctx, kf_c0)
    a = kf_c0.mask.copy()
    b = a.copy()
    b[:, : b.shape[1] // 2] = 0
    assert not np.array_equal(a, b)
    gpu_ctx.random_initialization(p)
    gpu_ctx.red_black_pass(p, 0, 1)
    got = []
    for m in (a, b, a):
        gpu_ctx.finalize(p, m)
        got.append(gpu_ctx.get_maps(2)[1].copy())
    want = []
    for k, m in enumerate((a, b)):
        gpu_ctx.image_put(9100 + k, m)
        gpu_ctx.set_mask(9100 + k)
        gpu_ctx.finalize(p)
        want.append(gpu_ctx.get_maps(2)[1].copy())
    gpu_ctx.set_mask(-1)
    for k in range(2):
        gpu_ctx.image_drop(9100 + k)
    assert not np.array_equal(want[0], want[1])
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[2], want[0])
    assert np.array_equal(got[1], want[1])
