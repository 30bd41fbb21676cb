"""pmp_clean_noalloc: the Cleaner with its device memory in a file-scope
array (the call allocates nothing), against pmp_clean on the same inputs,
bit for bit, on two calls in a row (the second smaller, so the array keeps
the first call's contents), plus the size limit."""
import numpy as np
import pytest

from patchmatch_amd import synth
from patchmatch_amd._lib import PmpError

pytestmark = pytest.mark.gpu


def _maps(seed, n, W, H):
    rng = np.random.default_rng(seed)
    traj = synth.trajectory(seed, n, 0.15)
    scene = synth.make_scene(seed + 40)
    depths = []
    for pose in traj:
        _, d = synth.render(scene, pose, W, H)
        d = d.astype(np.float32) * rng.choice([1.0, 1.0, 1.0, 1.2], d.shape).astype(np.float32)
        d[rng.random(d.shape) < 0.05] = 0.0
        depths.append(d)
    return depths, traj


def test_clean_noalloc_equals_clean(gpu_ctx):
    for step, (seed, n, W, H) in enumerate([(1, 5, 480, 240), (2, 3, 200, 100)]):
        depths, traj = _maps(seed, n, W, H)
        print(f"call {step}: {n} maps {W}x{H}", flush=True)
        want = gpu_ctx.clean(depths, traj, 0.03, 3)
        got = gpu_ctx.clean_noalloc(depths, traj, 0.03, 3)
        same = got.view(np.uint32) == want.view(np.uint32)
        assert same.all(), f"call {step}: {np.count_nonzero(~same)} of {same.size} differ"
        assert np.count_nonzero(want) > 0.05 * want.size
    # beyond the array: (16 + 1) * 640 * 480 floats > 2^22
    big = [np.ones((480, 640), np.float32)] * 16
    pose = np.concatenate([np.eye(3).ravel(), np.zeros(3)])
    with pytest.raises(PmpError, match="4194304"):
        gpu_ctx.clean_noalloc(big, [pose] * 16, 0.03, 3)
