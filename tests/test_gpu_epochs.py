"""Pass epochs of the persistent propagation kernel (pmp_debug_epochs): every
tile of every red-black pass runs exactly once, after the 8 tiles around it
have finished the previous pass, under the graph-replayed pmp_densify, the
directly launched pmp_red_black_pass, every launch shape, and the batched
form's work queue over keyframes of different sizes.  Mode 2 starts
tile 0 one pass ahead from the host, and the count shows exactly that one
violation; no kernel waits on an epoch, so every launch ends normally.  The
maps are the same bits in every mode."""
import numpy as np
import pytest

from patchmatch_amd.params import config1_params

pytestmark = pytest.mark.gpu


def _load(ctx, kf):
    ctx.load_keyframe(kf.ref, kf.views, kf.ref_pose, kf.view_poses)


def _bits(maps):
    return [np.ascontiguousarray(m).view(np.uint32).copy() for m in maps[:3]]


def _tiles(kf, tile_rows):
    return -(-((kf.W + 1) // 2) // 32) * -(-kf.H // (2 * tile_rows))


@pytest.mark.parametrize("tile_rows", [8, 2])
def test_densify_epochs(gpu_ctx, kf_small6, tile_rows):
    p = config1_params(iterations=3, seed=8)
    _load(gpu_ctx, kf_small6)
    cfg = gpu_ctx.batch_get_config()
    print(f"[epochs] densify 320x160, tile_rows {tile_rows}", flush=True)
    try:
        gpu_ctx.batch_set_config(tile_rows=tile_rows, blocks_per_cu=cfg["blocks_per_cu"], use_graph=1,
                                 xcd_bands=cfg["xcd_bands"])
        gpu_ctx.debug_epochs(0)
        gpu_ctx.densify(p)
        want = _bits(gpu_ctx.get_maps(2))
        for mode, bad in ((1, 0), (2, 1)):
            gpu_ctx.debug_epochs(mode)
            gpu_ctx.densify(p)
            rep = gpu_ctx.debug_epoch_report()
            got = _bits(gpu_ctx.get_maps(2))
            assert rep == {"tiles": _tiles(kf_small6, tile_rows), "violations": bad, "min_epoch": 6, "max_epoch": 6}, \
                (mode, rep)
            assert all(np.array_equal(a, b) for a, b in zip(got, want)), mode
    finally:
        gpu_ctx.debug_epochs(0)
        gpu_ctx.batch_set_config(**cfg)


def test_red_black_pass_epochs(gpu_ctx, kf_small6):
    """Iterations 1..3 from pmp_red_black_pass: passes 2..5, each launched
    directly; every tile ends at epoch 6."""
    p = config1_params(iterations=3, seed=9)
    _load(gpu_ctx, kf_small6)
    print("[epochs] red-black passes 2..5", flush=True)
    try:
        gpu_ctx.debug_epochs(0)
        gpu_ctx.random_initialization(p)
        gpu_ctx.red_black_pass(p, 0, 1)
        gpu_ctx.debug_epochs(1)
        gpu_ctx.red_black_pass(p, 1, 3)
        rep = gpu_ctx.debug_epoch_report()
        cfg = gpu_ctx.batch_get_config()
        assert rep == {"tiles": _tiles(kf_small6, cfg["tile_rows"]), "violations": 0, "min_epoch": 6,
                       "max_epoch": 6}, rep
    finally:
        gpu_ctx.debug_epochs(0)


def test_batch_epochs(gpu_ctx):
    """pmp_batch_densify over keyframes of three sizes (one work queue of
    (keyframe, tile) items per pass): every tile of every keyframe at
    2 x iterations, no violation; mode 2 shows exactly one; the maps are the
    same bits as without the check."""
    from patchmatch_amd import synth
    from patchmatch_amd._lib import Keyframe
    from conftest import cached
    kfs = cached("epoch_kfs", lambda: [synth.make_keyframe(20 + k, W, H, n, baseline=(0.05, 0.5), spheres=4)
                                       for k, (W, H, n) in enumerate([(320, 160, 6), (192, 96, 3), (448, 224, 2)])])
    p = config1_params(iterations=2, seed=10)
    descs = []
    for k, kf in enumerate(kfs):
        base = 5000 + 100 * k
        gpu_ctx.image_put(base, kf.ref)
        for v, img in enumerate(kf.views):
            gpu_ctx.image_put(base + 1 + v, img)
        descs.append(Keyframe.make(base, kf.ref_pose, [base + 1 + v for v in range(len(kf.views))], kf.view_poses))
    rows = gpu_ctx.batch_get_config()["tile_rows"]
    tiles = sum(_tiles(kf, rows) for kf in kfs)
    print(f"[epochs] batch of 3, {tiles} tiles", flush=True)
    try:
        gpu_ctx.debug_epochs(0)
        gpu_ctx.batch_densify(p, descs)
        want = [_bits(gpu_ctx.batch_get_maps(k, (kf.H, kf.W))) for k, kf in enumerate(kfs)]
        for mode, bad in ((1, 0), (2, 1)):
            gpu_ctx.debug_epochs(mode)
            gpu_ctx.batch_densify(p, descs)
            rep = gpu_ctx.debug_epoch_report()
            assert rep == {"tiles": tiles, "violations": bad, "min_epoch": 4, "max_epoch": 4}, (mode, rep)
            for k, kf in enumerate(kfs):
                got = _bits(gpu_ctx.batch_get_maps(k, (kf.H, kf.W)))
                assert all(np.array_equal(a, b) for a, b in zip(got, want[k])), (mode, k)
    finally:
        gpu_ctx.debug_epochs(0)
