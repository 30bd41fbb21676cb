"""The densify thread's loop in C++ (include/pmp/frontend.hpp, run by
examples/frontend_stream) against the same steps taken one by one through the
C-ABI: parallax gate, depth range from the landmarks, 3-view window with
landmark seeding and the warp of the previous result, 5-deep Cleaner.  The
cleaned depth maps must be equal bit for bit and belong to the same
keyframes."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from patchmatch_amd import _lib, synth
from patchmatch_amd._lib import PMP_MAPS_FINAL
from patchmatch_amd.params import config0_params

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "patchmatch_amd", "examples", "frontend_stream")


def _stream(W, H, n):
    traj = list(synth.trajectory(11, n, 0.25))
    traj.insert(4, traj[3].copy())                   # a keyframe without parallax: the gate drops it
    scene = synth.make_scene(110)
    rng = np.random.default_rng(4)
    rendered = [synth.render(scene, pose, W, H) for pose in traj]
    imgs = [synth.photometric(r[0], rng) for r in rendered]
    # a global landmark set on the surface seen from keyframe 0; every keyframe observes all of it
    B = synth.bearings(W, H)
    r, c = rng.integers(4, H - 4, 400), rng.integers(4, W - 4, 400)
    Xc = rendered[0][1][r, c][:, None] * B[r, c]
    xyz = (Xc - traj[0][9:]) @ traj[0][:9].reshape(3, 3)
    ids = np.arange(5000, 5400, dtype=np.uint64)
    return traj, imgs, ids, xyz


@pytest.mark.parametrize("px,deg", [(0.0, 0.0), (1.0, 30.0)])
def test_frontend_stream_equals_steps(gpu_ctx, tmp_path, px, deg):
    if not os.path.exists(EXE):
        pytest.fail("frontend_stream not built (run __graft_entry__.build())")
    W, H, it, thr = 160, 80, 2, 0.03
    traj, imgs, ids, xyz = _stream(W, H, 9)
    mask = np.ones((H, W), np.uint8)
    mask[int(0.85 * H):] = 0
    base = config0_params(seed=21)
    p = base.copy(iterations=it, patch_half=5)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(np.array([W, H, len(traj), it], np.int32).tobytes())
        f.write(np.array([thr, px, deg], np.float32).tobytes())
        f.write(ctypes.string_at(ctypes.addressof(base), ctypes.sizeof(base)))
        for k, (pose, img) in enumerate(zip(traj, imgs)):
            f.write(np.array([100 + k], np.uint64).tobytes())
            f.write(np.asarray(pose, np.float64).tobytes())
            f.write(img.tobytes())
            f.write(mask.tobytes())
            f.write(np.array([len(ids)], np.int64).tobytes())
            f.write(ids.tobytes())
            f.write(np.ascontiguousarray(xyz, np.float64).tobytes())
    out = tmp_path / "out.bin"
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    data = out.read_bytes()
    accepted = list(np.frombuffer(data[:4 * len(traj)], np.int32))
    rest = data[4 * len(traj):]
    per = 8 + 4 * W * H
    results = [(int(np.frombuffer(rest[q * per:q * per + 8], np.uint64)[0]),
                np.frombuffer(rest[q * per + 8:(q + 1) * per], np.float32).reshape(H, W))
               for q in range(len(rest) // per)]

    # the same loop, step by step through the C-ABI
    want_acc, window, depth_results, want = [], [], [], []
    gpu_ctx.forget_result()
    prev = None
    for k, (pose, img) in enumerate(zip(traj, imgs)):
        if prev is not None:
            s = _lib.score_view(traj[prev], ids, xyz, pose, ids, xyz)
            if not s > 0.25:
                want_acc.append(0)
                continue
        want_acc.append(1)
        window.append(k)
        prev = k
        if len(window) >= 3:
            ref, nb = window[-2], [window[-3], window[-1]]
            lo, hi = _lib.depth_interval(traj[ref], xyz)
            pk = p.copy(min_depth=lo, max_depth=hi)
            gpu_ctx.load_keyframe(imgs[ref], [imgs[m] for m in nb], traj[ref], [traj[m] for m in nb])
            gpu_ctx.random_initialization(pk)
            gpu_ctx.preinit_landmarks(pk, xyz, mask)
            gpu_ctx.preinit_warp(pk, mask)
            gpu_ctx.red_black_pass(pk, 0, pk.iterations)
            gpu_ctx.finalize(pk, mask)
            fin = gpu_ctx.get_maps(PMP_MAPS_FINAL)
            depth_results.append((ref, fin[1], fin[0]))
            gpu_ctx.memorize_result()
            if len(depth_results) >= 5:                # a new depth map fills the 5-deep queue
                q = depth_results[-5:]
                order = [2, 1, 0, 3, 4]                # swapCenter of the Cleaner (depthmap.cc:1301-1309)
                if px > 0 or deg > 0:
                    cleaned = gpu_ctx.clean_ex([q[o][1] for o in order], [q[o][2] for o in order],
                                               [traj[q[o][0]] for o in order], thr, 3, px, deg)
                else:
                    cleaned = gpu_ctx.clean([q[o][1] for o in order], [traj[q[o][0]] for o in order], thr, 3)
                want.append((100 + q[2][0], cleaned))
    assert accepted == want_acc and accepted[4] == 0
    print(f"[frontend] accepted {accepted}, cleaned keyframes {[kid for kid, _ in results]}", flush=True)
    assert [kid for kid, _ in results] == [kid for kid, _ in want] and len(want) >= 2
    for (kid, got), (_, exp) in zip(results, want):
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), kid
        if deg == 0:
            assert (got > 0).mean() > 0.05
