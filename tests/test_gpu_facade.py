"""The C++ facade (include/pmp/densifier.hpp) over a sliding 3-view window,
as the densify thread drives it (ComputePatchMatch with the landmark seeding
and the warp of the memorised previous result), equals direct C-ABI calls
bit for bit."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from patchmatch_amd import synth
from patchmatch_amd.params import config0_params

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "patchmatch_amd", "examples", "facade_window")


def test_facade_sliding_window(gpu_ctx, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("facade_window not built (run __graft_entry__.build())")
    W, H = 160, 80
    traj = synth.trajectory(7, 4, 0.25)
    scene = synth.make_scene(70)
    rng = np.random.default_rng(3)
    rendered = [synth.render(scene, pose, W, H) for pose in traj]
    imgs = [synth.photometric(r[0], rng) for r in rendered]
    B = synth.bearings(W, H)
    lms = []
    for pose, (_, gt) in zip(traj, rendered):       # landmarks on the surface seen by each keyframe
        r, c = rng.integers(6, H - 6, 150), rng.integers(6, W - 6, 150)
        Xc = gt[r, c][:, None] * B[r, c]
        lms.append((Xc - pose[9:]) @ pose[:9].reshape(3, 3))
    mask = np.ones((H, W), np.uint8)
    mask[int(0.85 * H):] = 0
    p = config0_params(iterations=2, seed=77, min_depth=0.05, max_depth=12.0)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(np.array([W, H, len(imgs)], np.int32).tobytes())
        f.write(ctypes.string_at(ctypes.addressof(p), ctypes.sizeof(p)))
        for pose, img, lm in zip(traj, imgs, lms):
            f.write(np.asarray(pose, np.float64).tobytes())
            f.write(img.tobytes())
            f.write(mask.tobytes())
            f.write(np.array([0.05, 12.0], np.float64).tobytes())
            f.write(np.array([len(lm)], np.int64).tobytes())
            f.write(np.ascontiguousarray(lm, np.float64).tobytes())
    out = tmp_path / "out.bin"
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    data = np.fromfile(out, dtype=np.uint8)
    P = W * H
    per = 3 * 4 * P
    assert data.size == 2 * per                        # windows [0,1,2] and [1,2,3]
    windows = [(1, [0, 2]), (2, [1, 3])]               # centre = reference, swapped order of the rest
    gpu_ctx.forget_result()
    accepted = []
    for w, (ref, nb) in enumerate(windows):
        chunk = data[w * per:(w + 1) * per]
        fd = chunk[:4 * P].view(np.float32).reshape(H, W)
        fs = chunk[4 * P:8 * P].view(np.float32).reshape(H, W)
        fn = chunk[8 * P:].view(np.int32).reshape(H, W)
        gpu_ctx.load_keyframe(imgs[ref], [imgs[k] for k in nb], traj[ref], traj[nb])
        gpu_ctx.random_initialization(p)
        accepted.append((gpu_ctx.preinit_landmarks(p, lms[ref], mask), gpu_ctx.preinit_warp(p, mask)))
        gpu_ctx.red_black_pass(p, 0, p.iterations)
        gpu_ctx.finalize(p, mask)
        _, gd, gs, gn = gpu_ctx.get_maps(2)
        gpu_ctx.memorize_result()
        assert np.array_equal(fd.view(np.uint32), gd.view(np.uint32))
        assert np.array_equal(fs.view(np.uint32), gs.view(np.uint32))
        assert np.array_equal(fn, gn)
        assert (gd > 0).mean() > 0.2
    print(f"[facade] pre-init accepted (landmarks, warp) per window: {accepted}", flush=True)
    assert accepted[0][0] > 0 and accepted[0][1] == 0 and accepted[1][1] > 0
