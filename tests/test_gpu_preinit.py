"""Pre-initialisation on the GPU (pmp_preinit_landmarks, pmp_memorize_result,
pmp_preinit_warp) against the oracle restatement, bit for bit: working maps,
copy-back snapshot and the number of accepted pixels."""
import numpy as np
import pytest

from oracle import or_py
from patchmatch_amd import synth
from patchmatch_amd._lib import PMP_MAPS_FINAL, PMP_MAPS_INIT, PMP_MAPS_WORK
from patchmatch_amd.params import config0_params, config1_params

pytestmark = pytest.mark.gpu


def _same(a, b, tag):
    eq = np.ascontiguousarray(a, np.float32).view(np.uint32) == np.ascontiguousarray(b, np.float32).view(np.uint32)
    assert eq.all(), f"{tag}: {np.count_nonzero(~eq)} of {eq.size} differ"


def _landmarks(kf, rng, n):
    """GT-surface points (distance jittered), duplicates on one pixel, and outliers."""
    B = synth.bearings(kf.W, kf.H)
    r = rng.integers(8, kf.H - 8, n)
    c = rng.integers(8, kf.W - 8, n)
    d = kf.gt_depth[r, c] * rng.uniform(0.95, 1.05, n)
    Xc = d[:, None] * B[r, c]
    Xc = np.concatenate([Xc, Xc[: n // 10] * 1.01, rng.uniform(-3, 3, (n // 20, 3))])
    R = kf.ref_pose[:9].reshape(3, 3)
    return (Xc - kf.ref_pose[9:]) @ R


@pytest.mark.parametrize("cfg", ["c0", "c1_topk"])
def test_preinit_landmarks_bitwise(gpu_ctx, kf_c0, cfg):
    p = config0_params() if cfg == "c0" else config1_params(iterations=2)
    okf = or_py.Keyframe.from_synth(kf_c0)
    work = or_py.init(okf, p)
    init = tuple(a.copy() for a in work)
    xyz = _landmarks(kf_c0, np.random.default_rng(7), 3000)
    w2, i2, _, acc = or_py.preinit_landmarks(okf, p, work, init, kf_c0.mask, xyz)
    print(f"[preinit landmarks {cfg}] {len(xyz)} landmarks, {acc} accepted", flush=True)
    gpu_ctx.load_keyframe(kf_c0.ref, kf_c0.views, kf_c0.ref_pose, kf_c0.view_poses)
    gpu_ctx.random_initialization(p)
    acc_g = gpu_ctx.preinit_landmarks(p, xyz, kf_c0.mask)
    assert acc_g == acc > 100
    for name, g, o in zip(("plane", "depth", "score"), gpu_ctx.get_maps(PMP_MAPS_WORK), w2):
        _same(g, o, f"work {name}")
    for name, g, o in zip(("plane", "depth", "score"), gpu_ctx.get_maps(PMP_MAPS_INIT), i2):
        _same(g, o, f"init {name}")


def test_preinit_warp_bitwise(gpu_ctx, kf_c0):
    """The previous result is a densification with neighbour 0 as the reference."""
    p = config0_params(iterations=2)
    kf = kf_c0
    # previous keyframe: view 0 as reference, (ref, view 1) as neighbours
    prev_pose = kf.view_poses[0]
    okf_prev = or_py.Keyframe(kf.views[0], [kf.ref, kf.views[1]], prev_pose, [kf.ref_pose, kf.view_poses[1]])
    prev = or_py.densify(okf_prev, p, kf.mask)
    gpu_ctx.image_put(0, kf.ref)
    gpu_ctx.image_put(1, kf.views[0])
    gpu_ctx.image_put(2, kf.views[1])
    gpu_ctx.set_views(1, prev_pose, [0, 2], [kf.ref_pose, kf.view_poses[1]])
    gpu_ctx.H, gpu_ctx.W = kf.H, kf.W
    gpu_ctx.densify(p, kf.mask)
    gp, gd, _, _ = gpu_ctx.get_maps(PMP_MAPS_FINAL)
    _same(gd, prev["depth"], "previous result depth")
    gpu_ctx.memorize_result()
    # current keyframe: warp the memorised result in
    okf = or_py.Keyframe.from_synth(kf)
    work = or_py.init(okf, p)
    init = tuple(a.copy() for a in work)
    prev4 = np.zeros((kf.H, kf.W, 4), np.float32)
    prev4[..., :3] = prev["plane"]
    w2, i2, _, acc = or_py.preinit_warp(okf, p, work, init, kf.mask, prev_pose, prev4, prev["depth"])
    print(f"[preinit warp] {acc} accepted", flush=True)
    gpu_ctx.set_views(0, kf.ref_pose, [1, 2], kf.view_poses)
    gpu_ctx.random_initialization(p)
    acc_g = gpu_ctx.preinit_warp(p, kf.mask)
    assert acc_g == acc > 100
    for name, g, o in zip(("plane", "depth", "score"), gpu_ctx.get_maps(PMP_MAPS_WORK), w2):
        _same(g, o, f"work {name}")
    for name, g, o in zip(("plane", "depth", "score"), gpu_ctx.get_maps(PMP_MAPS_INIT), i2):
        _same(g, o, f"init {name}")
    # and the passes continue from the pre-initialised maps exactly as the oracle's
    pl, dp, sc, _ = or_py.propagate(okf, p, *w2, 0, p.iterations)
    gpu_ctx.red_black_pass(p, 0, p.iterations)
    for name, g, o in zip(("plane", "depth", "score"), gpu_ctx.get_maps(PMP_MAPS_WORK), (pl, dp, sc)):
        _same(g, o, f"after passes {name}")
