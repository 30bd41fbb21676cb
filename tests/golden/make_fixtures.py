"""Writes tests/golden/tiny_densify.npz: inputs and oracle outputs of one tiny
keyframe (96x48, 2 neighbours, 7x7 patch, 2 iterations).  The oracle must keep
reproducing it bit for bit (a change of the canonical arithmetic shows up
here), and the GPU must match it.

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import or_py  # noqa: E402
from patchmatch_amd import synth  # noqa: E402
from patchmatch_amd.params import config0_params  # noqa: E402

PARAMS = dict(iterations=2, seed=1234)


def make():
    kf = synth.make_keyframe(11, 96, 48, 2, baseline=(0.1, 0.3))
    p = config0_params(**PARAMS)
    out = or_py.densify(or_py.Keyframe.from_synth(kf), p, kf.mask)
    return dict(ref=kf.ref, view0=kf.views[0], view1=kf.views[1], ref_pose=kf.ref_pose, view_poses=kf.view_poses,
                mask=kf.mask, **{k: v for k, v in out.items()})


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiny_densify.npz")
    np.savez_compressed(path, **make())
    print("wrote", path)
