"""Oracle pins of the host pre-initialisation (PreInitSparsePCL2,
PreInitResultWarp; depthmap.cc:361-516), restated in oracle/or_core.c.

What fixes the restatement independently of its own formulas:
  * landmarks placed on the ground-truth surface along a pixel's bearing land
    on that pixel, and the seeded plane has the landmark's distance as depth
    (1/|b.p| = distance) and the camera-facing normal -b;
  * several landmarks on one pixel resolve as the reference's sequential loop
    does: strictly better than the running score wins, so the first candidate
    with the highest score (re-checked here in Python from the candidate
    scores);
  * warping a result into the same camera returns every plane to its own
    pixel unchanged (n / (n.(d b)) with n = p/|p| and d = 1/(p.b) is p).
"""
import numpy as np
import pytest

from oracle import or_py
from patchmatch_amd import synth
from patchmatch_amd.params import config0_params

from conftest import gt_planes


def _setup(kf, p):
    okf = or_py.Keyframe.from_synth(kf)
    plane, depth, score = or_py.init(okf, p)
    return okf, (plane, depth, score), (plane.copy(), depth.copy(), score.copy())


def _landmarks(kf, pix, scale=1.0):
    """World points along the bearings of pixels pix (rows, cols) at scale x GT depth."""
    B = synth.bearings(kf.W, kf.H)
    r, c = pix
    Xc = (kf.gt_depth[r, c] * scale)[:, None] * B[r, c]
    R = kf.ref_pose[:9].reshape(3, 3)
    t = kf.ref_pose[9:]
    return (Xc - t) @ R          # R^T (Xc - t), row-wise


def test_landmarks_seed_their_pixel(kf_c0):
    p = config0_params()
    okf, work, init = _setup(kf_c0, p)
    rng = np.random.default_rng(3)
    rows = rng.integers(20, 200, 400)
    cols = rng.integers(20, 440, 400)
    key = np.unique(rows * kf_c0.W + cols)
    rows, cols = key // kf_c0.W, key % kf_c0.W
    xyz = _landmarks(kf_c0, (rows, cols))
    w2, i2, sc, acc = or_py.preinit_landmarks(okf, p, work, init, None, xyz)
    plane, depth, score = w2
    B = synth.bearings(kf_c0.W, kf_c0.H).astype(np.float32)
    scored = ~np.isnan(sc)
    assert scored.mean() > 0.95
    won = scored & (sc > work[2][rows, cols])
    assert acc == won.sum() > 0          # one landmark per pixel: pixels changed = winners
    r, c = rows[won], cols[won]
    np.testing.assert_allclose(depth[r, c], kf_c0.gt_depth[r, c], rtol=1e-5)
    pl = plane[r, c, :3]
    inv_d = np.abs(np.sum(pl * B[r, c], -1))
    np.testing.assert_allclose(1.0 / inv_d, depth[r, c], rtol=2e-5)
    # normal -b: the plane n / (n.(d b)) = b / d points along the bearing
    nrm = pl / np.linalg.norm(pl, axis=-1, keepdims=True)
    assert np.all(np.sum(nrm * B[r, c], -1) > 1 - 1e-5)
    # the snapshot for the copy-back holds the same values; untouched pixels unchanged
    for a, b in zip(w2, i2):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    changed = np.zeros((kf_c0.H, kf_c0.W), bool)
    changed[r, c] = True
    np.testing.assert_array_equal(depth[~changed].view(np.uint32), work[1][~changed].view(np.uint32))


def test_landmark_conflicts_resolve_sequentially(kf_c0):
    p = config0_params()
    okf, work, init = _setup(kf_c0, p)
    rng = np.random.default_rng(5)
    rows = rng.integers(30, 190, 60)
    cols = rng.integers(30, 430, 60)
    key = np.unique(rows * kf_c0.W + cols)
    rows, cols = key // kf_c0.W, key % kf_c0.W
    scales = [0.97, 1.0, 1.03, 1.0]       # four landmarks per pixel, one repeated (a tie)
    xyz = np.concatenate([_landmarks(kf_c0, (rows, cols), s) for s in scales])
    w2, _, sc, _ = or_py.preinit_landmarks(okf, p, work, init, None, xyz)
    n = len(rows)
    for q in range(n):
        k = rows[q] * kf_c0.W + cols[q]
        cur = work[2].reshape(-1)[k]
        win = None
        for m in range(len(scales)):          # the reference's loop order
            s = sc[m * n + q]
            if not np.isnan(s) and s > cur:
                cur, win = s, m
        got = w2[1].reshape(-1)[k]
        if win is None:
            assert got == work[1].reshape(-1)[k]
        else:
            assert got == np.float32(kf_c0.gt_depth.reshape(-1)[k] * scales[win]) or \
                np.isclose(got, kf_c0.gt_depth.reshape(-1)[k] * scales[win], rtol=1e-6)
            assert w2[2].reshape(-1)[k] == cur


def test_warp_into_the_same_camera_is_identity(kf_c0):
    p = config0_params()
    okf, work, init = _setup(kf_c0, p)
    prev = gt_planes(kf_c0)
    B = synth.bearings(kf_c0.W, kf_c0.H).astype(np.float32)
    pdepth = (1.0 / np.maximum(np.abs(np.sum(prev[..., :3] * B, -1)), 1e-6)).astype(np.float32)
    low = tuple(a.copy() for a in work)
    low[2][:] = -2.0                           # every candidate beats this
    w2, _, sc, acc = or_py.preinit_warp(okf, p, low, init, None, kf_c0.ref_pose, prev, pdepth)
    scored = ~np.isnan(sc)
    assert scored.mean() > 0.85
    moved = w2[2] > -2.0
    assert acc == moved.sum()
    # each accepted plane is its own pixel's previous plane
    r, c = np.nonzero(moved & scored)
    assert len(r) > 0.8 * scored.sum()
    np.testing.assert_allclose(w2[0][r, c, :3], prev[r, c, :3], rtol=2e-4, atol=1e-6)
    np.testing.assert_allclose(w2[1][r, c], pdepth[r, c], rtol=2e-5)


def test_warp_respects_masks_and_depth_floor(kf_c0):
    p = config0_params()
    okf, work, init = _setup(kf_c0, p)
    prev = gt_planes(kf_c0)
    pdepth = kf_c0.gt_depth.astype(np.float32).copy()
    pdepth[:, : kf_c0.W // 2] = 0.0             # invalid previous result: no candidates
    low = tuple(a.copy() for a in work)
    low[2][:] = -2.0
    mask = np.ones((kf_c0.H, kf_c0.W), np.uint8)
    mask[100:140] = 0
    w2, _, sc, _ = or_py.preinit_warp(okf, p, low, init, mask, kf_c0.ref_pose, prev, pdepth)
    assert np.all(np.isnan(sc[:, : kf_c0.W // 2]))
    assert np.all(np.isnan(sc[100:140]))
    assert np.all(w2[2][100:140] == -2.0)
