"""The synthetic keyframes are the same bits in every run: make_keyframe
renders large keyframes on worker threads, and a worker count or a run must
not change the images or the ground truth (the renderer's products are
elementwise numpy, since this image's BLAS gives varying results when
several threads call it at once)."""
import hashlib

import numpy as np

from patchmatch_amd import synth


def _digest(kf):
    h = hashlib.blake2b(digest_size=16)
    for a in [kf.ref, *kf.views, kf.mask, kf.gt_depth, kf.view_poses]:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_threaded_render_is_reproducible():
    # the benchmark keyframe (1920 x 960, 7 cameras on 8 threads); with BLAS
    # products in the renderer, three runs gave three different digests
    runs = [_digest(synth.config1(0, W=1920, H=960, n_views=6, workers=8)) for _ in range(3)]
    assert len(set(runs)) == 1, runs
