"""Batched and served workloads of the benchmark (synthetic, seeded).

Both draw keyframes from seeded trajectories through the config-1 scene
(room + 20 spheres + 1.2 m tube), one trajectory per resolution class,
rendered on the GPU (synth.sequence_gpu).  A keyframe's neighbours are its nearest
trajectory neighbours (0.2 m spacing), so consecutive requests share most of
their images, as a SLAM keyframe stream does.

batched: n keyframes with resolutions drawn from `classes` and view counts
    U{2..10}; depth maps/s of the batched form (one work-queue launch per
    stage over all keyframes, hipGraph, autotuned launch shape) against the
    plain form (pmp_densify per keyframe, images pooled) on the same set.
served: a Poisson stream of requests at a given rate; the served form
    (pmp_server, continuous batching, LRU pool) against the plain
    per-request path (upload the request's images, pmp_densify, read the
    maps back; one request at a time in arrival order) on the same stream.
    Latency is measured from a request's arrival to its maps on the host.
"""
from __future__ import annotations

import sys
import threading
import time

import numpy as np

BATCH_CLASSES = [(960, 480, 0.3), (1920, 960, 0.5), (2880, 1440, 0.2)]
SERVE_CLASSES = [(960, 480, 0.3), (1920, 960, 0.4), (2880, 1440, 0.2), (5760, 2880, 0.1)]


class Stream:
    """Rendered trajectories per resolution class and request specs over them."""

    def __init__(self, seed: int, classes, per_class: int = 16, spacing: float = 0.2, with_gt: bool = True):
        from . import synth
        self.classes = classes
        self.seqs = []
        for c, (W, H, _) in enumerate(classes):
            self.seqs.append(synth.sequence_gpu(seed * 31 + c, per_class, W, H, spacing, with_gt=with_gt))
            print(f"[workloads] rendered {per_class} keyframes at {W}x{H}", file=sys.stderr, flush=True)
        self.per_class = per_class

    def image_id(self, c, k):
        return 1_000_000 * (c + 1) + k

    def mask_id(self, c):
        return 1_000_000 * (c + 1) + 999_999

    def draw(self, rng, n, views=(2, 10)):
        """n request specs: (class, ref index, neighbour indices), with
        U{views[0]..views[1]} nearest neighbours."""
        w = np.array([c[2] for c in self.classes], dtype=np.float64)
        cls = rng.choice(len(self.classes), size=n, p=w / w.sum())
        out = []
        for c in cls:
            k = int(rng.integers(0, self.per_class))
            nv = int(rng.integers(views[0], views[1] + 1))
            others = sorted((m for m in range(self.per_class) if m != k), key=lambda m: (abs(m - k), m))
            out.append((int(c), k, others[:nv]))
        return out

    def keyframe(self, spec):
        c, k, nb = spec
        s = self.seqs[c]
        return dict(ref_id=self.image_id(c, k), ref_pose=s["poses"][k], ref=s["images"][k],
                    view_ids=[self.image_id(c, m) for m in nb], view_poses=[s["poses"][m] for m in nb],
                    views=[s["images"][m] for m in nb], mask_id=self.mask_id(c), mask=s["mask"],
                    gt=s["gt"][k], shape=s["images"][k].shape)


def digest(maps) -> str:
    """Hash of a keyframe's final maps (plane, depth, score, nghbr) bit for bit."""
    import hashlib
    h = hashlib.blake2b(digest_size=16)
    for a in maps[:4]:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def plain_digests(ctx, kfs, params):
    """The maps of each keyframe densified alone (pmp_densify, images pooled), hashed."""
    from ._lib import PMP_MAPS_FINAL
    out = []
    for kf in kfs:
        ctx.set_views(kf["ref_id"], kf["ref_pose"], kf["view_ids"], kf["view_poses"])
        ctx.set_mask(kf["mask_id"])
        ctx.densify(params)
        out.append(digest(ctx.get_maps(PMP_MAPS_FINAL)))
    return out


def accuracy(depth, gt, mask):
    unmasked = mask > 0
    valid = unmasked & (depth > 0)
    rel = np.abs(depth.astype(np.float64) - gt) / np.maximum(gt, 1e-9)
    return dict(valid_fraction=float(valid.sum() / max(unmasked.sum(), 1)),
                absrel_le_1pct_of_valid=float(np.mean(rel[valid] <= 0.01)) if valid.any() else 0.0,
                median_absrel=float(np.median(rel[valid])) if valid.any() else None)


def run_batched(stream: Stream, params, n: int, seed: int):
    from ._lib import PMP_MAPS_FINAL, Context, Keyframe
    rng = np.random.default_rng(seed)
    specs = stream.draw(rng, n)
    ctx = Context(0)
    descs, kfs = [], []
    put = set()
    for sp in specs:
        kf = stream.keyframe(sp)
        for iid, img in [(kf["ref_id"], kf["ref"])] + list(zip(kf["view_ids"], kf["views"])) + \
                [(kf["mask_id"], kf["mask"])]:
            if iid not in put:
                ctx.image_put(iid, img)
                put.add(iid)
        kfs.append(kf)
        descs.append(Keyframe.make(kf["ref_id"], kf["ref_pose"], kf["view_ids"], kf["view_poses"],
                                   mask_id=kf["mask_id"]))
    pix = sum(kf["shape"][0] * kf["shape"][1] for kf in kfs)
    # plain: one keyframe per call, images already pooled
    ctx.set_views(kfs[0]["ref_id"], kfs[0]["ref_pose"], kfs[0]["view_ids"], kfs[0]["view_poses"])
    ctx.densify(params)
    ctx.synchronize()
    t = time.perf_counter()
    for kf in kfs:
        ctx.set_views(kf["ref_id"], kf["ref_pose"], kf["view_ids"], kf["view_poses"])
        ctx.set_mask(kf["mask_id"])
        ctx.densify(params)
    ctx.synchronize()
    t_plain = time.perf_counter() - t
    ctx.H, ctx.W = kfs[-1]["shape"]
    last_plain = ctx.get_maps(PMP_MAPS_FINAL)
    # batched: autotune the launch shape on this batch, then time a densification
    cfg = ctx.batch_autotune(params, descs)
    ctx.batch_densify(params, descs)
    ctx.synchronize()
    t = time.perf_counter()
    ctx.batch_densify(params, descs)
    ctx.synchronize()
    t_batch = time.perf_counter() - t
    # every keyframe of the batch against the same keyframe densified alone (after the timing)
    acc, batch_dig = [], []
    for k in range(n):
        maps = ctx.batch_get_maps(k, kfs[k]["shape"])
        batch_dig.append(digest(maps))
        acc.append(accuracy(maps[1], kfs[k]["gt"], kfs[k]["mask"]))
    plain_dig = plain_digests(ctx, kfs, params)
    identical = sum(a == b for a, b in zip(plain_dig, batch_dig))
    same = batch_dig[-1] == digest(last_plain)
    ctx.close()
    res = {}
    for W, H, _ in stream.classes:
        res[f"{W}x{H}"] = sum(1 for kf in kfs if kf["shape"] == (H, W))
    return {
        "keyframes": n, "megapixels": round(pix / 1e6, 2), "resolutions": res,
        "views_mean": round(float(np.mean([len(kf["view_ids"]) for kf in kfs])), 2),
        "plain_depth_maps_per_s": round(n / t_plain, 3), "batched_depth_maps_per_s": round(n / t_batch, 3),
        "speedup": round(t_plain / t_batch, 3), "plain_s": round(t_plain, 3), "batched_s": round(t_batch, 3),
        "autotuned_config": cfg, "last_keyframe_bit_identical": bool(same),
        "keyframes_bit_identical_to_plain": int(identical),
        "accuracy_mean": {k: round(float(np.mean([a[k] for a in acc if a[k] is not None])), 4)
                          for k in ("valid_fraction", "absrel_le_1pct_of_valid", "median_absrel")},
    }


def run_pipeline(seed: int = 0, n: int = 30, W: int = 1920, H: int = 960, views=(4, 8), params=None,
                 thr: float = 0.01, max_reproj_px: float = 1.0, max_normal_deg: float = 10.0, keep: bool = False,
                 base_freq: float | None = None, warm_iterations=(4, 2), warm_views: int | None = None):
    """The whole dense path over n keyframes of one trajectory through the
    config-1 scene (0.2 m spacing), as the reference's two threads run it:
      densify  every keyframe against its U{views} nearest trajectory
               neighbours (one batched call);
      clean    each map against its two predecessors and two successors
               (DepthmapCleaner: 5-deep queue, centre map first, at least 3
               consistent views, depth within thr);
      fuse     the cleaned maps in order (DepthmapPruner: 5-deep queue, prune
               the head, pop it; dense_pointcloud_module.cc:56-182).
    The same windows are also cleaned with pmp_clean_ex's round trip within
    max_reproj_px, normals within max_normal_deg, and both.  Then the
    keyframes are densified again one by one in trajectory order with the
    densify thread's warm start (ComputePatchMatch with PreInitResultWarp of
    the previous keyframe's memorised result, depthmap.cc:434-516, 808-829):
    random init, warp of the previous result, the red-black passes, copy-back,
    memorise; at the params' iteration count and at each of warm_iterations
    (default 4 and 2: the reference's example and code defaults,
    config.cc:137-190), each with the
    keyframes' latency (all but the first, which has nothing to warp).  Accuracy against the rendered ground truth after each stage;
    with keep, the maps of every stage are returned too (for parity checks).
    base_freq: the texture scale (default: the config-1 scene's).
    warm_views: when given, the warm-start runs densify every keyframe against
    exactly this many nearest trajectory neighbours instead of U{views}."""
    from . import synth
    from ._lib import Context, Keyframe
    from .params import config1_params
    p = params if params is not None else config1_params(seed=seed)
    seq = synth.sequence_gpu(seed, n, W, H, 0.2,
                             base_freq=synth.TEX_FREQ_C1 if base_freq is None else base_freq)
    poses, imgs, gts, mask = seq["poses"], seq["images"], seq["gt"], seq["mask"]
    rng = np.random.default_rng(seed + 0x2C)
    ctx = Context(0)
    t0 = time.perf_counter()
    for k in range(n):
        ctx.image_put(k + 1, imgs[k])
    ctx.image_put(n + 1, mask)
    descs = []
    for k in range(n):
        nv = int(rng.integers(views[0], views[1] + 1))
        nb = sorted((m for m in range(n) if m != k), key=lambda m: (abs(m - k), m))[:nv]
        descs.append(Keyframe.make(k + 1, poses[k], [m + 1 for m in nb], [poses[m] for m in nb], mask_id=n + 1))
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.batch_densify(p, descs)
    maps = [ctx.batch_get_maps(k, (H, W)) for k in range(n)]       # plane4, depth, score, nghbr
    t2 = time.perf_counter()
    order = [2, 1, 0, 3, 4]                                          # swapCenter: centre map first
    centres = list(range(2, n - 2))
    windows = {c: [c - 2 + o for o in order] for c in centres}
    cleaned = {c: ctx.clean([maps[m][1] for m in windows[c]], [poses[m] for m in windows[c]], thr, 3)
               for c in centres}
    t3 = time.perf_counter()
    ctx.pruner_reset()
    fused, pruned_maps = [], {}
    q = 0
    for c in centres:
        ctx.pruner_add_view(cleaned[c], poses[c], np.repeat(imgs[c][..., None], 3, -1))
        if ctx.pruner_size() == 5:
            h = centres[q]
            pr, pts, cols, kps = ctx.pruner_prune(thr, (H, W))
            ctx.pruner_pop_head()
            fused.append((h, pts, kps))
            pruned_maps[h] = pr
            q += 1
    t4 = time.perf_counter()
    checked = {}
    for name, px, deg in (("round_trip", max_reproj_px, 0.0), ("normals", 0.0, max_normal_deg),
                          ("round_trip_and_normals", max_reproj_px, max_normal_deg)):
        checked[name] = {c: ctx.clean_ex([maps[m][1] for m in windows[c]], [maps[m][0] for m in windows[c]],
                                         [poses[m] for m in windows[c]], thr, 3, px, deg) for c in centres}
    t5 = time.perf_counter()
    # warm start: each keyframe pre-initialised from the previous one's result,
    # at the params' iteration count and at the reference's defaults of 4
    # (example config) and 2 (code default); latency per keyframe from the
    # views' selection to the final maps on the device
    from ._lib import PMP_MAPS_FINAL

    def warm_run(iterations):
        ctx.forget_result()
        maps_w, warped, lat = {}, [], []
        for k in range(n):
            if warm_views:
                nb = [m + 1 for m in sorted((m for m in range(n) if m != k), key=lambda m: (abs(m - k), m))[:warm_views]]
            else:
                nb = [int(descs[k].view_ids[v]) for v in range(descs[k].n_views)]
            ts = time.perf_counter()
            ctx.set_views(k + 1, poses[k], nb, [poses[m - 1] for m in nb])
            ctx.random_initialization(p)
            if k > 0:
                warped.append(ctx.preinit_warp(p, mask))
            ctx.red_black_pass(p, 0, iterations)
            ctx.finalize(p, mask)
            ctx.memorize_result()
            ctx.synchronize()
            lat.append((time.perf_counter() - ts) * 1e3)
            maps_w[k] = ctx.get_maps(PMP_MAPS_FINAL)[1]
        return maps_w, warped, lat

    warm_iters = sorted({p.iterations, *warm_iterations}, reverse=True)
    warm_runs = {it: warm_run(it) for it in warm_iters}
    warm = warm_runs[p.iterations][0]
    t6 = time.perf_counter()
    ctx.close()

    unm = mask > 0

    def summary(dmaps):
        nv = n1 = 0
        rels = []
        for c, d in dmaps.items():
            valid = (d > 0) & unm
            rel = np.abs(d.astype(np.float64) - gts[c]) / np.maximum(gts[c], 1e-9)
            nv += valid.sum()
            n1 += np.sum(rel[valid] <= 0.01)
            rels.append(rel[valid])
        r = np.concatenate(rels) if rels else np.zeros(0)
        return {"valid_fraction": round(float(nv / (unm.sum() * len(dmaps))), 4),
                "absrel_le_1pct_of_valid": round(float(n1 / max(nv, 1)), 4),
                "median_absrel": round(float(np.median(r)), 5) if r.size else None}

    # fused points: distance to the head keyframe's camera against the ground truth at their pixel
    prel, npts = [], 0
    for h, pts, kps in fused:
        npts += len(pts)
        if len(pts):
            R, t = poses[h][:9].reshape(3, 3), poses[h][9:]
            dist = np.linalg.norm(pts @ R.T + t, axis=1)
            gt = gts[h][kps[:, 0], kps[:, 1]]
            prel.append(np.abs(dist - gt) / gt)
    prel = np.concatenate(prel) if prel else np.zeros(0)
    acc = {"densified": summary({c: maps[c][1] for c in centres}), "cleaned": summary(cleaned)}
    for it, (mw, warped, lat) in warm_runs.items():
        key = "densified_warm_start" if it == p.iterations else f"densified_warm_start_{it}_iterations"
        acc[key] = summary({c: mw[c] for c in centres})
        acc[key]["warped_candidates_accepted_mean"] = round(float(np.mean(warped)), 1) if warped else 0.0
        acc[key]["latency_ms"] = {"p50": round(float(np.percentile(lat[1:], 50)), 3),
                                  "p99": round(float(np.percentile(lat[1:], 99)), 3)}
    acc.update({f"cleaned_{k}": summary(v) for k, v in checked.items()})
    acc["fused_points"] = {"count": int(npts), "prunes": len(fused),
                           "absrel_le_1pct": round(float(np.mean(prel <= 0.01)), 4) if prel.size else None,
                           "median_absrel": round(float(np.median(prel)), 5) if prel.size else None}
    out = {
        "keyframes": n, "width": W, "height": H, "views": list(views), "same_depth_thr": thr,
        "max_reproj_px": max_reproj_px, "max_normal_deg": max_normal_deg,
        "seconds": {"upload": round(t1 - t0, 3), "densify_batched": round(t2 - t1, 3), "clean": round(t3 - t2, 3),
                    "fuse": round(t4 - t3, 3), "clean_with_checks_x3": round(t5 - t4, 3),
                    "densify_warm_start_one_by_one": round(t6 - t5, 3)},
        "warm_start_iterations": warm_iters, "warm_start_views": warm_views or list(views),
        "keyframes_per_s": round(n / (t4 - t1), 3),
        "accuracy": acc,
    }
    if keep:
        out.update(_maps=maps, _cleaned=cleaned, _checked=checked, _windows=windows, _poses=poses, _images=imgs,
                   _fused=fused, _pruned=pruned_maps, _warm=warm)
    return out


def _arrivals(rng, rate, n):
    return np.cumsum(rng.exponential(1.0 / rate, n))


def _summary(lat, t_first, t_last, n, labels=None):
    """Throughput and latency percentiles; with labels (one per request, e.g.
    its resolution), the percentiles per label too."""
    lat = np.asarray(lat) * 1e3
    out = {"throughput_per_s": round(n / (t_last - t_first), 3), "p50_ms": round(float(np.percentile(lat, 50)), 1),
           "p99_ms": round(float(np.percentile(lat, 99)), 1), "max_ms": round(float(lat.max()), 1)}
    if labels is not None:
        by = {}
        for lab in sorted(set(labels)):
            v = lat[np.asarray(labels) == lab]
            by[lab] = {"n": int(v.size), "p50_ms": round(float(np.percentile(v, 50)), 1),
                       "p99_ms": round(float(np.percentile(v, 99)), 1)}
        out["by_resolution"] = by
    return out


def _labels(stream, specs):
    return [f"{stream.classes[c][0]}x{stream.classes[c][1]}" for c, _, _ in specs]


def run_plain_stream(stream: Stream, params, specs, arrivals):
    """One request at a time in arrival order: upload its images, densify, read back."""
    from ._lib import PMP_MAPS_FINAL, Context
    ctx = Context(0)
    # warm the kernels once (not timed)
    kf = stream.keyframe(specs[0])
    ctx.load_keyframe(kf["ref"], kf["views"], kf["ref_pose"], kf["view_poses"])
    ctx.densify(params, kf["mask"])
    ctx.synchronize()
    lat, done = [], []
    t0 = time.perf_counter()
    for spec, ta in zip(specs, arrivals):
        now = time.perf_counter() - t0
        if now < ta:
            time.sleep(ta - now)
        kf = stream.keyframe(spec)
        ctx.load_keyframe(kf["ref"], kf["views"], kf["ref_pose"], kf["view_poses"])
        ctx.densify(params, kf["mask"])
        ctx.get_maps(PMP_MAPS_FINAL)
        t = time.perf_counter() - t0
        lat.append(t - ta)
        done.append(t)
    ctx.close()
    return _summary(lat, arrivals[0], max(done), len(specs), _labels(stream, specs))


def run_served_stream(stream: Stream, params, specs, arrivals, max_slots=16, hbm_budget=64 << 30, inflight_kpx=0,
                      tail_stages=0):
    """Requests submitted at their arrival times to a pmp_server."""
    from ._lib import Server
    srv = Server(0, max_slots=max_slots, hbm_budget=hbm_budget, edge_mode=params.edge_mode, inflight_kpx=inflight_kpx,
                 tail_stages=tail_stages)
    # warm the kernels once (not timed)
    kf = stream.keyframe(specs[0])
    t = srv.submit(params, kf["ref_id"], kf["ref_pose"], kf["view_ids"], kf["view_poses"], kf["shape"],
                   ref=kf["ref"], views=kf["views"], mask_id=kf["mask_id"], mask=kf["mask"])
    srv.wait(t)
    srv.fetch(t)
    tickets = [None] * len(specs)
    t_py0 = time.perf_counter()
    base = {}

    def producer():
        for q, (spec, ta) in enumerate(zip(specs, arrivals)):
            now = time.perf_counter() - t_py0
            if now < ta:
                time.sleep(ta - now)
            kf = stream.keyframe(spec)
            tickets[q] = srv.submit(params, kf["ref_id"], kf["ref_pose"], kf["view_ids"], kf["view_poses"],
                                    kf["shape"], ref=kf["ref"], views=kf["views"], mask_id=kf["mask_id"],
                                    mask=kf["mask"])
            base.setdefault("first", time.perf_counter())

    th = threading.Thread(target=producer)
    th.start()
    lat, done = [], []
    infos = []
    for q in range(len(specs)):
        while tickets[q] is None:
            time.sleep(0.0005)
        srv.wait(tickets[q])
        *_, info = srv.fetch(tickets[q])
        infos.append(info)
    th.join()
    # server clock -> this clock: the warm-up request's fetch anchors nothing; use the
    # first submission's own submit stamp against the time the producer recorded
    t_srv0 = infos[0]["submit_ms"] / 1e3
    off = base["first"] - t_py0 - t_srv0      # py time (relative to t_py0) = srv time + off
    for info, ta in zip(infos, arrivals):
        d = info["done_ms"] / 1e3 + off
        lat.append(d - ta)
        done.append(d)
    st = srv.stats()
    srv.close()
    out = _summary(lat, arrivals[0], max(done), len(specs), _labels(stream, specs))
    out.update({"max_in_flight": st["max_in_flight"], "uploads": st["uploads"], "evictions": st["evictions"],
                "steps": st["steps"]})
    return out


def run_served(stream: Stream, params, rates, n_requests: int, seed: int, views=(2, 10), below=0.5, tail_stages=0):
    """The same request list at every rate; only the arrival times change.
    First every request arrives at once: the completion rate of that
    saturated run is each form's capacity.  Then the given rates, and
    `below` x the served capacity, where arrivals do not outrun the device
    and the latencies are those of the requests rather than of a backlog.
    With tail_stages > 0, the below-capacity rate is also served with that
    latency-first admission (pmp_server_opts.tail_stages)."""
    out = {}
    specs = stream.draw(np.random.default_rng(seed), n_requests, views)
    at_once = np.zeros(n_requests)
    cap_plain = run_plain_stream(stream, params, specs, at_once)
    cap_served = run_served_stream(stream, params, specs, at_once)
    out["capacity"] = {"plain_per_s": cap_plain["throughput_per_s"], "served_per_s": cap_served["throughput_per_s"],
                       "ratio": round(cap_served["throughput_per_s"] / cap_plain["throughput_per_s"], 3)}
    print(f"[workloads] capacity: {out['capacity']}", file=sys.stderr, flush=True)
    rates = list(rates) + ([round(below * cap_served["throughput_per_s"], 2)] if below else [])
    for rate in rates:
        arr = _arrivals(np.random.default_rng(seed + int(rate * 1000)), rate, n_requests)
        plain = run_plain_stream(stream, params, specs, arr)
        print(f"[workloads] {rate:g} req/s plain: {plain}", file=sys.stderr, flush=True)
        served = run_served_stream(stream, params, specs, arr)
        print(f"[workloads] {rate:g} req/s served: {served}", file=sys.stderr, flush=True)
        out[f"{rate:g}_per_s"] = {"plain": plain, "served": served,
                                  "throughput_ratio": round(served["throughput_per_s"] / plain["throughput_per_s"], 3),
                                  "p99_ratio": round(served["p99_ms"] / plain["p99_ms"], 3)}
        if tail_stages and below and rate == rates[-1]:
            lf = run_served_stream(stream, params, specs, arr, tail_stages=tail_stages)
            print(f"[workloads] {rate:g} req/s served, tail_stages {tail_stages}: {lf}", file=sys.stderr, flush=True)
            out[f"{rate:g}_per_s"]["served_latency_first"] = dict(
                lf, tail_stages=tail_stages, p50_ratio=round(lf["p50_ms"] / plain["p50_ms"], 3),
                p99_ratio=round(lf["p99_ms"] / plain["p99_ms"], 3))
    return out


def image_bytes(stream: Stream, specs):
    """Pool bytes the requests' distinct images need (1 B per pixel, 3 B once
    packed as a neighbour), and the largest need of one request."""
    need, packed, one = {}, set(), 0
    for c, k, nb in specs:
        W, H = stream.classes[c][0], stream.classes[c][1]
        ids = [(stream.image_id(c, k), False), (stream.mask_id(c), False)] + \
            [(stream.image_id(c, m), True) for m in nb]
        one = max(one, sum(W * H * (3 if pk else 1) for _, pk in ids))
        for iid, pk in ids:
            need[iid] = W * H
            if pk:
                packed.add(iid)
    return sum(b * (3 if iid in packed else 1) for iid, b in need.items()), one


def verify_served(stream: Stream, params, specs, max_slots=16, hbm_budget=64 << 30, window=None):
    """Every request of `specs`, submitted in order to a pmp_server with the
    given slots and HBM budget and at most `window` requests outstanding
    (default max_slots + 1), against the same keyframe densified alone
    (pmp_densify): the number of bit-identical results, and the server's
    evictions and uploads."""
    from ._lib import Context, Server
    ctx = Context(0)
    kfs = [stream.keyframe(sp) for sp in specs]
    want, put = [], set()
    for kf in kfs:
        for iid, img in [(kf["ref_id"], kf["ref"])] + list(zip(kf["view_ids"], kf["views"])) + \
                [(kf["mask_id"], kf["mask"])]:
            if iid not in put:
                ctx.image_put(iid, img)
                put.add(iid)
        want.extend(plain_digests(ctx, [kf], params))
    ctx.close()
    srv = Server(0, max_slots=max_slots, hbm_budget=hbm_budget, edge_mode=params.edge_mode)
    window = window or max_slots + 1
    tickets, got = [], []
    for q, kf in enumerate(kfs):
        if len(tickets) - len(got) >= window:
            srv.wait(tickets[len(got)])
            got.append(digest(srv.fetch(tickets[len(got)])))
        tickets.append(srv.submit(params, kf["ref_id"], kf["ref_pose"], kf["view_ids"], kf["view_poses"],
                                  kf["shape"], ref=kf["ref"], views=kf["views"], mask_id=kf["mask_id"],
                                  mask=kf["mask"]))
    while len(got) < len(tickets):
        srv.wait(tickets[len(got)])
        got.append(digest(srv.fetch(tickets[len(got)])))
    st = srv.stats()
    srv.close()
    return {"requests": len(specs), "bit_identical_to_plain": int(sum(a == b for a, b in zip(want, got))),
            "evictions": st["evictions"], "uploads": st["uploads"], "max_in_flight": st["max_in_flight"],
            "hbm_budget": hbm_budget}
