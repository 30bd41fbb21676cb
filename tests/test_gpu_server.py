"""Served form (pmp_server_*): requests of mixed sizes, view counts and
params densified with continuous batching.  Each request's maps must be
bit-identical to the one-keyframe path alone (pmp_densify) -- whether the
server runs it alone (one slot) or in the mix with others, submitted at once
or while others are on the device -- and the image pool must evict by LRU
within its HBM budget."""
import threading
import time

import numpy as np
import pytest

from patchmatch_amd import synth
from patchmatch_amd._lib import PMP_MAPS_FINAL, PmpError, Server
from patchmatch_amd.params import EDGE_CLAMP, config1_params, default_params

from conftest import cached

pytestmark = pytest.mark.gpu

SHAPES = [(15, 320, 160, 6), (16, 256, 128, 2), (17, 384, 192, 4), (18, 192, 96, 10), (19, 224, 112, 5),
          (20, 192, 96, 1), (21, 320, 160, 3)]


def _keyframes():
    def make():
        return [synth.make_keyframe(s, W, H, n, baseline=(0.05, 0.5), spheres=8, tube=True) for s, W, H, n in SHAPES]
    return cached("server_kfs", make)


def _params(k):
    # every request with its own params: aggregation, iterations, seed, depth range
    if k % 3 == 0:
        return config1_params(iterations=2, seed=100 + k)
    if k % 3 == 1:
        return default_params(iterations=1, seed=100 + k)
    return config1_params(iterations=3, seed=100 + k, topk=2, max_depth=9.0)


def _ids(k, n):
    base = 1000 * (k + 1)
    return base, [base + 1 + v for v in range(n)], base + 999


def _plain(gpu_ctx, kfs):
    def run():
        out = []
        for k, kf in enumerate(kfs):
            ref, views, mask = _ids(k, len(kf.views))
            gpu_ctx.image_put(ref, kf.ref)
            for v, img in enumerate(kf.views):
                gpu_ctx.image_put(views[v], img)
            gpu_ctx.set_views(ref, kf.ref_pose, views, kf.view_poses)
            gpu_ctx.H, gpu_ctx.W = kf.H, kf.W
            gpu_ctx.densify(_params(k), kf.mask)
            out.append(gpu_ctx.get_maps(PMP_MAPS_FINAL))
        return out
    return cached("server_plain", run)


def _submit(srv, kfs, k, with_pixels=True):
    kf = kfs[k]
    ref, views, mask = _ids(k, len(kf.views))
    return srv.submit(_params(k), ref, kf.ref_pose, views, kf.view_poses, (kf.H, kf.W),
                      ref=kf.ref if with_pixels else None, views=kf.views if with_pixels else None,
                      mask_id=mask, mask=kf.mask if with_pixels else None)


def _check(got, want, tag):
    for name, u, v in zip(("plane", "depth", "score", "nghbr"), got[:4], want):
        same = np.ascontiguousarray(u).view(np.uint32) == np.ascontiguousarray(v).view(np.uint32)
        assert same.all(), f"{tag}: {name}: {np.count_nonzero(~same)} of {same.size} differ"


@pytest.mark.parametrize("slots", [1, 4, 16])
def test_server_same_alone_and_in_the_mix(gpu_ctx, slots):
    kfs = _keyframes()
    want = _plain(gpu_ctx, kfs)
    print(f"[server slots={slots}] {len(kfs)} requests submitted at once", flush=True)
    with Server(0, max_slots=slots) as srv:
        srv.pause(True)          # queue everything, then admit as many as fit at once
        tickets = [_submit(srv, kfs, k) for k in range(len(kfs))]
        assert srv.stats()["queued"] == len(kfs)
        srv.pause(False)
        for k, t in enumerate(tickets):
            assert srv.wait(t, 60000), f"request {k} not done in 60 s"
            got = srv.fetch(t)
            _check(got, want[k], f"slots={slots} request {k}")
            assert got[4]["steps"] == 2 * _params(k).iterations + 2
        st = srv.stats()
        assert st["completed"] == len(kfs) and st["failed"] == 0
        assert st["max_in_flight"] == min(slots, len(kfs))


def test_server_latency_first_admission(gpu_ctx):
    """tail_stages = 2: requests submitted one after another (each while the
    previous is in flight) join it only for its last 2 stages, unless a backlog
    builds up; every result is unchanged."""
    kfs = _keyframes()
    want = _plain(gpu_ctx, kfs)
    print("[server] latency-first admission, requests one after another", flush=True)
    with Server(0, max_slots=4, tail_stages=2) as srv:
        tickets = []
        for k in range(len(kfs)):
            tickets.append(_submit(srv, kfs, k))
            time.sleep(0.002)
        for k, t in enumerate(tickets):
            assert srv.wait(t, 60000), f"request {k} not done in 60 s"
            _check(srv.fetch(t), want[k], f"latency-first request {k}")
        st = srv.stats()
        assert st["completed"] == len(kfs) and st["failed"] == 0


def test_server_inflight_pixel_budget(gpu_ctx):
    """An admission budget of 120 thousand pixels: requests are admitted oldest
    first while the keyframes in flight stay within it (a larger one alone),
    results unchanged."""
    kfs = _keyframes()
    want = _plain(gpu_ctx, kfs)
    budget_px = 120_000
    with Server(0, max_slots=16, inflight_kpx=budget_px // 1000) as srv:
        srv.pause(True)
        tickets = [_submit(srv, kfs, k) for k in range(len(kfs))]
        srv.pause(False)
        infos = []
        for k, t in enumerate(tickets):
            assert srv.wait(t, 60000), f"request {k} not done in 60 s"
            got = srv.fetch(t)
            _check(got, want[k], f"budget request {k}")
            infos.append(got[4])
        st = srv.stats()
        print(f"[server budget] {st}", flush=True)
        assert st["completed"] == len(kfs) and 1 < st["max_in_flight"] < len(kfs)
    # admission is oldest first: admit times are non-decreasing in submission order
    adm = [i["admit_ms"] for i in infos]
    assert adm == sorted(adm)
    # at every admission, the keyframes in flight (admitted, not yet done) fit the budget
    for k, i in enumerate(infos):
        flying = [j for j in range(len(infos)) if infos[j]["admit_ms"] <= i["admit_ms"] < infos[j]["done_ms"]]
        px = sum(kfs[j].W * kfs[j].H for j in flying)
        assert len(flying) == 1 or px <= budget_px, (k, flying, px)


def test_server_landmark_seeding(gpu_ctx):
    """A request with landmarks equals the one-keyframe sequence init ->
    pmp_preinit_landmarks -> passes -> finalize, in the mix with a request
    without landmarks; the seeding changes the result."""
    kfs = _keyframes()
    want_plain = _plain(gpu_ctx, kfs)
    k = 0
    kf = kfs[k]
    p = _params(k)
    rng = np.random.default_rng(5)
    B = synth.bearings(kf.W, kf.H)
    r, c = rng.integers(8, kf.H - 8, 600), rng.integers(8, kf.W - 8, 600)
    Xc = (kf.gt_depth[r, c] * rng.uniform(0.97, 1.03, 600))[:, None] * B[r, c]
    Xc = np.concatenate([Xc, rng.uniform(-3, 3, (40, 3))])
    xyz = (Xc - kf.ref_pose[9:]) @ kf.ref_pose[:9].reshape(3, 3)
    ref, views, mask = _ids(k, len(kf.views))
    gpu_ctx.image_put(ref, kf.ref)
    for v, img in enumerate(kf.views):
        gpu_ctx.image_put(views[v], img)
    gpu_ctx.set_views(ref, kf.ref_pose, views, kf.view_poses)
    gpu_ctx.H, gpu_ctx.W = kf.H, kf.W
    gpu_ctx.random_initialization(p)
    acc = gpu_ctx.preinit_landmarks(p, xyz, kf.mask)
    gpu_ctx.red_black_pass(p, 0, p.iterations)
    gpu_ctx.finalize(p, kf.mask)
    want = gpu_ctx.get_maps(PMP_MAPS_FINAL)
    print(f"[server landmarks] {len(xyz)} landmarks, {acc} accepted on the plain path", flush=True)
    assert acc > 50
    assert not np.array_equal(want[1].view(np.uint32), want_plain[k][1].view(np.uint32))
    with Server(0, max_slots=4) as srv:
        srv.pause(True)
        t0 = srv.submit(p, ref, kf.ref_pose, views, kf.view_poses, (kf.H, kf.W), ref=kf.ref, views=kf.views,
                        mask_id=mask, mask=kf.mask, landmarks=xyz)
        t1 = _submit(srv, kfs, 1)
        srv.pause(False)
        assert srv.wait(t0, 60000) and srv.wait(t1, 60000)
        _check(srv.fetch(t0), want, "seeded request")
        _check(srv.fetch(t1), want_plain[1], "request without landmarks beside it")


def test_server_warm_start(gpu_ctx):
    """A request with the previous keyframe's result (warm start) equals the
    one-keyframe sequence densify(previous) -> pmp_memorize_result ->
    init -> pmp_preinit_warp -> passes -> finalize, in the mix with a request
    without one; the warm start changes the result."""
    kfs = _keyframes()
    want_plain = _plain(gpu_ctx, kfs)
    seq = synth.sequence_gpu(21, 4, 320, 160, with_gt=False, base_freq=synth.TEX_FREQ_C0)
    imgs, poses, mask = seq["images"], seq["poses"], seq["mask"]
    ids = [9000 + k for k in range(4)]
    for k in range(4):
        gpu_ctx.image_put(ids[k], imgs[k])
    p = config1_params(iterations=2, seed=77)
    nb0, nb1 = [1, 2, 3], [0, 2, 3]
    gpu_ctx.set_views(ids[0], poses[0], [ids[m] for m in nb0], [poses[m] for m in nb0])
    gpu_ctx.densify(p, mask)
    prev = gpu_ctx.get_maps(PMP_MAPS_FINAL)
    gpu_ctx.memorize_result()
    gpu_ctx.set_views(ids[1], poses[1], [ids[m] for m in nb1], [poses[m] for m in nb1])
    gpu_ctx.random_initialization(p)
    acc = gpu_ctx.preinit_warp(p, mask)
    gpu_ctx.red_black_pass(p, 0, p.iterations)
    gpu_ctx.finalize(p, mask)
    want = gpu_ctx.get_maps(PMP_MAPS_FINAL)
    gpu_ctx.forget_result()
    gpu_ctx.densify(p, mask)
    cold = gpu_ctx.get_maps(PMP_MAPS_FINAL)
    print(f"[server warm start] {acc} warped candidates accepted on the plain path", flush=True)
    assert acc > 1000
    assert not np.array_equal(want[1].view(np.uint32), cold[1].view(np.uint32))
    with Server(0, max_slots=4) as srv:
        srv.pause(True)
        t0 = srv.submit(p, ids[1], poses[1], [ids[m] for m in nb1], [poses[m] for m in nb1], mask.shape,
                        ref=imgs[1], views=[imgs[m] for m in nb1], mask_id=9999, mask=mask,
                        prev=(prev[0], prev[1], poses[0]))
        t1 = _submit(srv, kfs, 1)
        srv.pause(False)
        assert srv.wait(t0, 60000) and srv.wait(t1, 60000)
        _check(srv.fetch(t0), want, "warm-started request")
        _check(srv.fetch(t1), want_plain[1], "request without a warm start beside it")
        with pytest.raises(ValueError):
            srv.submit(p, ids[1], poses[1], [ids[m] for m in nb1], [poses[m] for m in nb1], mask.shape,
                       mask_id=9999, prev=(prev[0][:-1], prev[1], poses[0]))


def test_server_staggered_arrivals(gpu_ctx):
    """Requests submitted from another thread while earlier ones run."""
    kfs = _keyframes()
    want = _plain(gpu_ctx, kfs)
    with Server(0, max_slots=3) as srv:
        tickets = {}

        def producer():
            for k in range(len(kfs)):
                tickets[k] = _submit(srv, kfs, k)
                time.sleep(0.003 * (k % 3))
        th = threading.Thread(target=producer)
        th.start()
        th.join(60)
        assert len(tickets) == len(kfs)
        for k in range(len(kfs)):
            assert srv.wait(tickets[k], 60000)
            _check(srv.fetch(tickets[k]), want[k], f"staggered request {k}")
        print(f"[server staggered] {srv.stats()}", flush=True)


def test_server_pool_lru_and_budget(gpu_ctx):
    """A budget that holds the maps of one slot plus a few images: images are
    evicted least-recently-used first, pooled ids are reused without pixels,
    and evicted ids must be given pixels again."""
    kfs = _keyframes()
    want = _plain(gpu_ctx, kfs)
    k0 = 0   # 320x160, 6 views
    kf = kfs[k0]
    P = kf.W * kf.H
    budget = P * 84 + P * (1 + 3 * 6 + 1) + 3 * P * 3     # maps; grey + 2-byte packed column pairs per view
    with Server(0, max_slots=1, hbm_budget=budget) as srv:
        t = _submit(srv, kfs, k0)
        assert srv.wait(t, 60000)
        _check(srv.fetch(t), want[k0], "first")
        # pooled: no pixels needed, no uploads
        up = srv.stats()["uploads"]
        t = _submit(srv, kfs, k0, with_pixels=False)
        assert srv.wait(t, 60000)
        _check(srv.fetch(t), want[k0], "pooled")
        assert srv.stats()["uploads"] == up
        # another keyframe of the same size does not fit next to the first: evictions
        k6 = 6   # 320x160, 3 views
        t = _submit(srv, kfs, k6)
        assert srv.wait(t, 60000)
        _check(srv.fetch(t), want[k6], "after eviction")
        st = srv.stats()
        assert st["evictions"] > 0 and st["pool_bytes"] + st["maps_bytes"] <= budget
        # the first keyframe's reference was evicted (oldest): pixels are required again
        with pytest.raises(PmpError, match="ENOTFOUND"):
            _submit(srv, kfs, k0, with_pixels=False)


def test_server_rejects_bad_requests(gpu_ctx):
    kfs = _keyframes()
    with Server(0, max_slots=2) as srv:
        kf = kfs[1]
        ref, views, mask = _ids(1, len(kf.views))
        with pytest.raises(PmpError, match="EINVAL"):
            srv.submit(_params(1).copy(edge_mode=EDGE_CLAMP), ref, kf.ref_pose, views, kf.view_poses, (kf.H, kf.W),
                       ref=kf.ref, views=kf.views)
        with pytest.raises(PmpError, match="ENOMEM"):
            Server(0, max_slots=1, hbm_budget=1000).submit(_params(1), ref, kf.ref_pose, views, kf.view_poses,
                                                          (kf.H, kf.W), ref=kf.ref, views=kf.views)
        with pytest.raises(PmpError, match="ENOTFOUND"):
            srv.wait(987654, 0)


def test_server_cancel_and_reset(gpu_ctx):
    """pmp_server_cancel ends a queued request at once and one in flight once
    its issued steps have run; pmp_server_reset cancels everything queued or
    in flight and empties the pool.  The requests that ran are bit-identical
    to the plain path, and the server keeps serving after both."""
    kfs = _keyframes()
    want = _plain(gpu_ctx, kfs)
    with Server(0, max_slots=2) as srv:
        srv.pause(True)
        t = [_submit(srv, kfs, k) for k in range(len(kfs))]
        srv.cancel(t[3])          # queued
        srv.pause(False)
        srv.cancel(t[0])          # admitted first: in flight (or already finished)
        done = {}
        for k in range(len(kfs)):
            assert srv.wait(t[k], 60000)
            try:
                done[k] = srv.fetch(t[k])
            except PmpError as e:
                assert "ECANCELED" in str(e) and k in (0, 3), (k, str(e))
        assert 3 not in done and len(done) >= len(kfs) - 2
        for k, got in done.items():
            _check(got, want[k], f"request {k} next to cancelled ones")
        assert srv.stats()["cancelled"] >= 1
        # reset with requests queued and in flight (their images are pooled)
        t2 = [_submit(srv, kfs, k, with_pixels=False) for k in (1, 2, 4, 5, 6)]
        srv.reset()
        st = srv.stats()
        print(f"[server reset] {st}", flush=True)
        assert st["in_flight"] == 0 and st["queued"] == 0 and st["pool_images"] == 0 and st["pool_bytes"] == 0
        for k, tk in zip((1, 2, 4, 5, 6), t2):
            assert srv.wait(tk, 0)
            try:
                _check(srv.fetch(tk), want[k], f"request {k} finished before the reset")
            except PmpError as e:
                assert "ECANCELED" in str(e)
        with pytest.raises(PmpError, match="ENOTFOUND"):
            _submit(srv, kfs, 1, with_pixels=False)
        tk = _submit(srv, kfs, 1)
        assert srv.wait(tk, 60000)
        _check(srv.fetch(tk), want[1], "after the reset")
