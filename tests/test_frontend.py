"""Keyframe gate and depth range of the densify thread (pmp_score_view,
pmp_depth_interval; host code, no GPU) against float64 restatements of
depthmap_computation_module.cc:123-198 and data/keyframe.cc:428-451, plus
geometries whose answer is known in closed form."""
import math

import numpy as np

from patchmatch_amd import _lib


def _pose(centre, yaw=0.0):
    c, s = math.cos(yaw), math.sin(yaw)
    R = np.array([[c, 0, -s], [0, 1, 0], [s, 0, c]])          # world->camera
    t = -R @ np.asarray(centre, float)
    return np.concatenate([R.ravel(), t])


def _ref_score(prev_pose, cur_pose, lms):
    def centre(p):
        R = p[:9].reshape(3, 3)
        return -R.T @ p[9:]
    cp, cc = centre(prev_pose), centre(cur_pose)
    score = 0.0
    for lm in lms:
        a, b = cc - lm, cp - lm
        theta = math.acos(a @ b / math.sqrt((a @ a) * (b @ b)))
        score += 1.0 if math.pi / 60 < theta < math.pi / 6 else 0.0
    return score / len(lms) if len(lms) else float("nan")


def test_score_view_matches_restatement_and_matches_by_id():
    rng = np.random.default_rng(0)
    xyz = rng.uniform(-4, 4, (300, 3))
    ids = np.arange(1000, 1300, dtype=np.uint64)
    prev, cur = _pose([0, 0, 0], 0.1), _pose([0.4, 0.05, -0.2], -0.3)
    sel_p, sel_c = np.arange(0, 200), np.arange(100, 300)          # 100 common landmarks
    got = _lib.score_view(prev, ids[sel_p], xyz[sel_p], cur, ids[sel_c], xyz[sel_c])
    want = _ref_score(prev, cur, xyz[100:200])
    assert got == want and 0 < got < 1


def test_score_view_closed_forms():
    lm = np.array([[0.0, 0.0, 5.0]])
    ids = np.array([7], np.uint64)
    # baseline b at distance d: angle 2 atan(b / 2d)
    for b, inside in [(0.1, False), (1.0, True), (5.0, False)]:     # 1.1 deg, 11.4 deg, 53 deg
        got = _lib.score_view(_pose([-b / 2, 0, 0]), ids, lm, _pose([b / 2, 0, 0]), ids, lm)
        assert got == (1.0 if inside else 0.0), b
    same = _lib.score_view(_pose([0, 0, 0]), ids, lm, _pose([0, 0, 0]), ids, lm)
    assert same == 0.0                                              # no parallax
    none = _lib.score_view(_pose([0, 0, 0]), ids, lm, _pose([1, 0, 0]), np.array([8], np.uint64), lm)
    assert math.isnan(none)                                         # 0/0: rejected by the > 0.25 gate


def test_depth_interval():
    rng = np.random.default_rng(1)
    pose = _pose([0.3, -0.1, 0.2], 0.7)
    for n in (1, 9, 10, 11, 57, 1000):
        xyz = rng.uniform(-5, 5, (n, 3))
        lo, hi = _lib.depth_interval(pose, xyz)
        R, t = pose[:9].reshape(3, 3), pose[9:]
        d = np.sort(np.linalg.norm(xyz @ R.T + t, axis=1).astype(np.float32))
        assert lo == 1e-5
        assert hi == float(d[n // 10 * 9]) * 1.1
    try:
        _lib.depth_interval(pose, np.zeros((0, 3)))
        raise AssertionError("empty landmark set accepted")
    except _lib.PmpError:
        pass
