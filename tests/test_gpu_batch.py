"""Batched form (pmp_batch_*): many keyframes of different sizes and view
counts per launch.  Every keyframe's final maps must be bit-identical to the
one-keyframe path (pmp_set_views + pmp_densify) with the same params, for
every launch configuration, with and without the captured hipGraph, and
across graph replays with new descriptors."""
import numpy as np
import pytest

from patchmatch_amd import synth
from patchmatch_amd._lib import PMP_MAPS_FINAL, Keyframe
from patchmatch_amd.params import AGG_MAX, EDGE_ZERO, config1_params, default_params

from conftest import cached

pytestmark = pytest.mark.gpu

# (seed, W, H, views): view counts cover every group split (3+3+3+1, 3+2, 3+1, 2, 1)
SHAPES = [(5, 320, 160, 6), (6, 256, 128, 2), (7, 384, 192, 4), (8, 192, 96, 10), (9, 224, 112, 5),
          (10, 192, 96, 1)]


def _keyframes():
    def make():
        return [synth.make_keyframe(s, W, H, n, baseline=(0.05, 0.5), spheres=8, tube=True) for s, W, H, n in SHAPES]
    return cached("batch_kfs", make)


def _upload(ctx, kfs, depth_override=None):
    out = []
    for k, kf in enumerate(kfs):
        base = 1000 * (k + 1)
        ctx.image_put(base, kf.ref)
        for v, img in enumerate(kf.views):
            ctx.image_put(base + 1 + v, img)
        ctx.image_put(base + 999, kf.mask)
        lo, hi = (0.0, 0.0) if depth_override is None else depth_override[k]
        out.append(Keyframe.make(base, kf.ref_pose, [base + 1 + v for v in range(len(kf.views))], kf.view_poses,
                                 mask_id=base + 999, min_depth=lo, max_depth=hi))
    return out


def _plain(ctx, kfs, p, depth_override=None):
    res = []
    for k, kf in enumerate(kfs):
        base = 1000 * (k + 1)
        ctx.set_views(base, kf.ref_pose, [base + 1 + v for v in range(len(kf.views))], kf.view_poses)
        ctx.H, ctx.W = kf.H, kf.W
        pk = p
        if depth_override is not None and depth_override[k][1] > 0:
            pk = p.copy(min_depth=depth_override[k][0], max_depth=depth_override[k][1])
        ctx.densify(pk, kf.mask)
        res.append(ctx.get_maps(PMP_MAPS_FINAL))
    return res


def _batch(ctx, kfs, descs, p):
    ctx.batch_densify(p, descs)
    return [ctx.batch_get_maps(k, (kf.H, kf.W)) for k, kf in enumerate(kfs)]


def _assert_same(a, b, tag):
    for k, (x, y) in enumerate(zip(a, b)):
        for name, u, v in zip(("plane", "depth", "score", "nghbr"), x, y):
            same = np.ascontiguousarray(u).view(np.uint32) == np.ascontiguousarray(v).view(np.uint32)
            assert same.all(), f"{tag}: keyframe {k} {name}: {np.count_nonzero(~same)} of {same.size} differ"


PARAMS = {
    "topk3_full": lambda: config1_params(iterations=2),
    "ref_fold_grid": lambda: default_params(iterations=2),
    "max_zero": lambda: default_params(iterations=1, agg_mode=AGG_MAX, edge_mode=EDGE_ZERO, seed=77),
}


@pytest.mark.parametrize("name", sorted(PARAMS))
def test_batch_equals_plain(gpu_ctx, name):
    kfs = _keyframes()
    p = PARAMS[name]()
    print(f"[batch {name}] {[(kf.W, kf.H, len(kf.views)) for kf in kfs]}", flush=True)
    over = [(0.0, 0.0)] * len(kfs)
    over[2] = (0.5, 6.0)
    descs = _upload(gpu_ctx, kfs, over)
    ref = _plain(gpu_ctx, kfs, p, over)
    got = _batch(gpu_ctx, kfs, descs, p)
    _assert_same(got, ref, name)


def test_batch_config_invariance_and_graph_replay(gpu_ctx):
    kfs = _keyframes()
    p = config1_params(iterations=2)
    descs = _upload(gpu_ctx, kfs)
    ref = _plain(gpu_ctx, kfs, p)
    for cfg in [(2, 1, 0, 0), (8, 2, 1, 1), (4, 0, 1, 0), (4, 3, 0, 1), (4, 0, 1, 1)]:
        print(f"[batch config] rows/bpc/graph/bands = {cfg}", flush=True)
        gpu_ctx.batch_set_config(*cfg)
        _assert_same(_batch(gpu_ctx, kfs, descs, p), ref, f"config {cfg}")
    # replay the cached graph with other descriptors: a reordered sub-batch
    order = [3, 0, 5, 1]
    sub = _batch(gpu_ctx, [kfs[k] for k in order], [descs[k] for k in order], p)
    _assert_same(sub, [ref[k] for k in order], "graph replay, reordered sub-batch")
    # and the full batch again
    _assert_same(_batch(gpu_ctx, kfs, descs, p), ref, "graph replay, full batch")


def test_batch_autotune_keeps_results(gpu_ctx):
    kfs = _keyframes()
    p = config1_params(iterations=1)
    descs = _upload(gpu_ctx, kfs)
    ref = _plain(gpu_ctx, kfs, p)
    best = gpu_ctx.batch_autotune(p, descs)
    print(f"[batch autotune] {best}", flush=True)
    assert best["tile_rows"] in (2, 4, 8) and best["blocks_per_cu"] >= 1
    assert gpu_ctx.batch_get_config() == best
    _assert_same(_batch(gpu_ctx, kfs, descs, p), ref, "after autotune")
    gpu_ctx.batch_set_config()


def test_batch_largest_size_with_small(gpu_ctx):
    """A 5760x2880 keyframe (the largest served size, 4 neighbours) batched
    with a 960x480 one equals each alone, bit for bit."""
    from patchmatch_amd import synth
    kfs = []
    for seed, (W, H) in ((23, (5760, 2880)), (24, (960, 480))):
        seq = synth.sequence_gpu(seed, 5, W, H, with_gt=False)
        ims, poses = seq["images"], seq["poses"]
        nb = (0, 1, 3, 4) if W == 5760 else (1, 3)
        kfs.append(synth.Keyframe(W=W, H=H, ref=ims[2], views=[ims[k] for k in nb], ref_pose=poses[2],
                                  view_poses=np.stack([poses[k] for k in nb]), mask=seq["mask"], gt_depth=None,
                                  bgr=None))
    p = config1_params(iterations=1, seed=9)
    descs = _upload(gpu_ctx, kfs)
    ref = _plain(gpu_ctx, kfs, p)
    _assert_same(_batch(gpu_ctx, kfs, descs, p), ref, "5760x2880 + 960x480")


def test_batch_rejects_bad_input(gpu_ctx):
    from patchmatch_amd._lib import PmpError
    kfs = _keyframes()
    descs = _upload(gpu_ctx, kfs)
    bad = Keyframe.make(424242, kfs[0].ref_pose, [1001], [kfs[0].view_poses[0]])
    with pytest.raises(PmpError, match="ENOTFOUND"):
        gpu_ctx.batch_densify(default_params(iterations=1), [descs[0], bad])
    mixed = Keyframe.make(1000, kfs[0].ref_pose, [2001], [kfs[0].view_poses[0]])   # 320x160 ref, 256x128 view
    with pytest.raises(PmpError, match="EINVAL"):
        gpu_ctx.batch_densify(default_params(iterations=1), [mixed])
    with pytest.raises(PmpError, match="EINVAL"):
        gpu_ctx.batch_get_maps(5, (160, 320))
