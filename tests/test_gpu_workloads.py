"""The batched and served forms at the benchmark's workloads, and against
the oracle directly.

* The batched mix: 64 keyframes of 960x480 / 1920x960 / 2880x1440 (weights
  0.3/0.5/0.2) with U{2..10} views from the seeded trajectories, in one
  pmp_batch_densify, under two launch configurations: every keyframe equals
  the same keyframe densified alone, bit for bit.
* The served stream: requests of up to 5760x2880 with U{2..10} views, under
  an HBM budget small enough to evict images: every request equals the same
  keyframe densified alone.
* Small batched keyframes and served requests against or_densify (the
  oracle) directly."""
import numpy as np
import pytest

from oracle import or_py
from patchmatch_amd import synth, workloads
from patchmatch_amd._lib import Context, Keyframe, Server
from patchmatch_amd.params import AGG_REF_FOLD, config0_params, config1_params

from conftest import cached

pytestmark = pytest.mark.gpu


def _batch_stream():
    return cached("c3_stream", lambda: workloads.Stream(11, workloads.BATCH_CLASSES))


@pytest.mark.parametrize("tile_rows,blocks_per_cu", [(8, 2), (4, 3)])
def test_batched_mix_every_keyframe_equals_plain(tile_rows, blocks_per_cu):
    st = _batch_stream()
    p = config1_params(seed=3)
    specs = st.draw(np.random.default_rng(12), 64)
    kfs = [st.keyframe(sp) for sp in specs]
    print(f"[batched mix] {len(kfs)} keyframes, {sum(kf['shape'][0] * kf['shape'][1] for kf in kfs) / 1e6:.1f} Mpx, "
          f"config ({tile_rows}, {blocks_per_cu})", flush=True)
    with Context(0) as ctx:
        put = set()
        for kf in kfs:
            for iid, img in [(kf["ref_id"], kf["ref"])] + list(zip(kf["view_ids"], kf["views"])) + \
                    [(kf["mask_id"], kf["mask"])]:
                if iid not in put:
                    ctx.image_put(iid, img)
                    put.add(iid)
        descs = [Keyframe.make(kf["ref_id"], kf["ref_pose"], kf["view_ids"], kf["view_poses"], mask_id=kf["mask_id"])
                 for kf in kfs]
        ctx.batch_set_config(tile_rows=tile_rows, blocks_per_cu=blocks_per_cu)
        ctx.batch_densify(p, descs)
        got = [workloads.digest(ctx.batch_get_maps(k, kfs[k]["shape"])) for k in range(len(kfs))]
        want = workloads.plain_digests(ctx, kfs, p)
    bad = [k for k in range(len(kfs)) if got[k] != want[k]]
    assert not bad, f"keyframes {bad} differ from pmp_densify"


def test_served_stream_with_evictions_equals_plain():
    st = cached("c4_stream", lambda: workloads.Stream(13, workloads.SERVE_CLASSES, per_class=16, with_gt=False))
    p = config1_params(seed=4)
    rng = np.random.default_rng(14)
    specs = st.draw(rng, 16)
    # four 5760x2880 requests over different images, 2 to 4 views; the first
    # two open the stream, so both slots grow to 5760x2880 at once
    specs[0], specs[1], specs[11], specs[14] = (3, 2, [1, 3, 4]), (3, 9, [8, 10, 7]), (3, 13, [12, 14, 15, 11]), \
        (3, 5, [4, 6])
    big = 5760 * 2880
    total, one = workloads.image_bytes(st, specs)
    # maps of two 5760x2880 slots plus the images of the two requests that
    # may be outstanding at once: once both slots have grown, the pool holds
    # at most `room`, below the four large requests' images alone, so pooled
    # images must be evicted
    room = 2 * one + (64 << 20)
    assert workloads.image_bytes(st, [specs[q] for q in (0, 1, 11, 14)])[0] > room
    assert total > 1.3 * room
    res = workloads.verify_served(st, p, specs, max_slots=2, hbm_budget=2 * 84 * big + room, window=2)
    print(f"[served stream] images {total / 1e9:.2f} GB, room {room / 1e9:.2f} GB: {res}", flush=True)
    assert res["bit_identical_to_plain"] == res["requests"]
    assert res["evictions"] > 0


def test_batched_and_served_equal_the_oracle():
    """Three small keyframes (the reference's pooled fold, 2-3 views) batched in
    one call and served as requests: each equals or_densify bit for bit."""
    kfs = [synth.make_keyframe(40 + k, W, H, n, baseline=(0.1, 0.3)) for k, (W, H, n) in
           enumerate([(160, 80, 2), (128, 64, 3), (192, 96, 2)])]
    p = config0_params(iterations=2, seed=9, coverage_mode=1)
    assert p.agg_mode == AGG_REF_FOLD
    want = [or_py.densify(or_py.Keyframe.from_synth(kf), p, kf.mask) for kf in kfs]
    with Context(0) as ctx:
        descs = []
        for k, kf in enumerate(kfs):
            base = 100 * (k + 1)
            ctx.image_put(base, kf.ref)
            for v, img in enumerate(kf.views):
                ctx.image_put(base + 1 + v, img)
            ctx.image_put(base + 99, kf.mask)
            descs.append(Keyframe.make(base, kf.ref_pose, [base + 1 + v for v in range(len(kf.views))],
                                       kf.view_poses, mask_id=base + 99))
        ctx.batch_densify(p, descs)
        for k, kf in enumerate(kfs):
            plane, depth, score, nghbr = ctx.batch_get_maps(k, (kf.H, kf.W))
            w = want[k]
            assert np.array_equal(depth.view(np.uint32), w["depth"].view(np.uint32)), f"batched {k} depth"
            assert np.array_equal(score.view(np.uint32), w["score"].view(np.uint32)), f"batched {k} score"
            assert np.array_equal(plane[..., :3].view(np.uint32), w["plane"].view(np.uint32)), f"batched {k} plane"
            assert np.array_equal(nghbr, w["nghbr"]), f"batched {k} nghbr"
    with Server(0, max_slots=2) as srv:
        tickets = []
        for k, kf in enumerate(kfs):
            base = 100 * (k + 1)
            tickets.append(srv.submit(p, base, kf.ref_pose, [base + 1 + v for v in range(len(kf.views))],
                                      kf.view_poses, (kf.H, kf.W), ref=kf.ref, views=kf.views, mask_id=base + 99,
                                      mask=kf.mask))
        for k, t in enumerate(tickets):
            assert srv.wait(t, 60000)
            plane, depth, score, nghbr, _ = srv.fetch(t)
            w = want[k]
            assert np.array_equal(depth.view(np.uint32), w["depth"].view(np.uint32)), f"served {k} depth"
            assert np.array_equal(score.view(np.uint32), w["score"].view(np.uint32)), f"served {k} score"
            assert np.array_equal(plane[..., :3].view(np.uint32), w["plane"].view(np.uint32)), f"served {k} plane"
            assert np.array_equal(nghbr, w["nghbr"]), f"served {k} nghbr"
