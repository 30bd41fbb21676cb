"""The whole dense path over one trajectory (densify -> reference Cleaner
-> Pruner, plus the Cleaner with the round-trip and normal tests), as
workloads.run_pipeline runs it at 1920x960 for the 30-keyframe sequence:
here 14 keyframes at 480x240 (with config 0's texture scale, which suits
that width).

Every stage is checked against the oracle on the stage's own inputs (the
first cleaning window with and without the tests, the first prune);
cleaning must raise the fraction of depths within 1 % of the rendered
ground truth."""
import numpy as np
import pytest

from oracle import or_py
from patchmatch_amd import synth, workloads
from patchmatch_amd.params import config1_params

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32 if a.dtype == np.float32 else np.uint64)


def test_pipeline_stages_match_oracle():
    n, W, H = 14, 480, 240
    p = config1_params(seed=3, iterations=4)
    r = workloads.run_pipeline(seed=3, n=n, W=W, H=H, params=p, keep=True, base_freq=synth.TEX_FREQ_C0)
    acc = r["accuracy"]
    print(f"[pipeline] {acc}", flush=True)
    maps, poses, imgs = r["_maps"], r["_poses"], r["_images"]
    c = 2
    win = r["_windows"][c]
    dm, pl = [maps[m][1] for m in win], [maps[m][0] for m in win]
    o_plain = or_py.clean(dm, poses[win], 0.01, 3)
    assert np.array_equal(_bits(o_plain), _bits(r["_cleaned"][c]))
    for name, px, deg in (("round_trip", 1.0, 0.0), ("normals", 0.0, 10.0), ("round_trip_and_normals", 1.0, 10.0)):
        o = or_py.clean_ex(dm, pl, poses[win], 0.01, 3, px, deg)
        assert np.array_equal(_bits(o), _bits(r["_checked"][name][c])), name
        assert not np.any((o > 0) & (o_plain == 0))
    # first prune: the head and the four later cleaned maps as they were queued
    heads = list(range(2, n - 2))
    depths = [r["_cleaned"][h].copy() for h in heads[:5]]
    bgr = np.repeat(imgs[heads[0]][..., None], 3, -1)
    pr, pts, cols, kps = or_py.prune(depths, poses[heads[:5]], bgr, 0.01)
    h0, gpts, gkps = r["_fused"][0]
    assert h0 == heads[0]
    assert np.array_equal(_bits(pr), _bits(r["_pruned"][h0]))
    assert np.array_equal(_bits(pts), _bits(gpts)) and np.array_equal(kps, gkps)
    # cleaning trades coverage for precision; fusion keeps points
    d, cl = acc["densified"], acc["cleaned"]
    assert cl["absrel_le_1pct_of_valid"] > d["absrel_le_1pct_of_valid"]
    assert cl["valid_fraction"] < d["valid_fraction"]
    assert acc["fused_points"]["count"] > 0
