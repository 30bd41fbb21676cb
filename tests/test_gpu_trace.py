"""roctx ranges around the library's calls (pmp_set_roctx): with tracing on,
the one-keyframe, batched and served paths give the same maps as with it off
(the ranges only bracket the host calls)."""
import numpy as np
import pytest

from patchmatch_amd import _lib, synth
from patchmatch_amd._lib import Context, Keyframe, Server
from patchmatch_amd.params import config1_params

from conftest import cached

pytestmark = pytest.mark.gpu


def _run(kf, p):
    out = []
    with Context(0) as ctx:
        ctx.load_keyframe(kf.ref, kf.views, kf.ref_pose, kf.view_poses, base_id=10)
        ctx.densify(p)
        out.append(ctx.get_maps(2)[:3])
        d = Keyframe.make(10, kf.ref_pose, [11 + v for v in range(len(kf.views))], kf.view_poses)
        ctx.batch_densify(p, [d])
        out.append(ctx.batch_get_maps(0, (kf.H, kf.W))[:3])
    with Server(0, max_slots=2) as srv:
        t = srv.submit(p, 10, kf.ref_pose, [11 + v for v in range(len(kf.views))], kf.view_poses, (kf.H, kf.W),
                       ref=kf.ref, views=kf.views)
        assert srv.wait(t, 60000)
        out.append(srv.fetch(t)[:3])
    return out


def test_ranges_leave_results_unchanged():
    kf = cached("trace_kf", lambda: synth.make_keyframe(31, 256, 128, 3, baseline=(0.05, 0.5), spheres=3))
    p = config1_params(iterations=2, seed=3)
    print("[trace] 256x128, 3 views, plain/batched/served", flush=True)
    want = _run(kf, p)
    try:
        assert _lib.set_roctx(True)
        got = _run(kf, p)
    finally:
        _lib.set_roctx(False)
    for w, g in zip(want, got):
        for a, b in zip(w, g):
            assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
