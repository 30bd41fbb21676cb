"""Workload driver for rocprofv3: densify the 1920x960 + 6-view keyframe
`--steps` times, optionally run the densify -> clean -> fuse trajectory of
`--pipeline` keyframes at the same size (the streaming Cleaner and Pruner
kernels), then run a streaming copy of known size (calibration of the
hardware byte counters: a copy of B bytes moves B bytes in and B bytes out).

    python -m patchmatch_amd.profile_run --steps 2 --copy-mib 256
    python -m patchmatch_amd.profile_run --steps 1 --pipeline 10 --copy-mib 256
    python -m patchmatch_amd.profile_run --steps 0 --batched 64 --copy-mib 0
    python -m patchmatch_amd.profile_run --steps 1 --params '{"edge_mode": "zero"}'
    python -m patchmatch_amd.profile_run --steps 0 --warm 12 --warm-iterations 1 --copy-mib 0

--batched n runs the batched workload of the benchmark (n keyframes of mixed
resolution and view count, plain then batched, workloads.run_batched).
--warm n densifies n keyframes of a trajectory (0.2 m apart) one by one with
the densify thread's warm start, each against its --views nearest neighbours
at --warm-iterations iterations, and prints each keyframe's latency from the
selection of its views to its final maps on the device.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def warm_sequence(ctx, p, n, W, H, views, iterations):
    import numpy as np
    from patchmatch_amd import synth
    seq = synth.sequence_gpu(5, n, W, H, 0.2, base_freq=synth.TEX_FREQ_C1)
    poses, imgs = seq["poses"], seq["images"]
    for k in range(n):
        ctx.image_put(2001 + k, imgs[k])
    ctx.image_put(2000, seq["mask"])
    ctx.set_timing(False)
    ctx.forget_result()
    lat = []
    for k in range(n):
        nb = sorted((m for m in range(n) if m != k), key=lambda m: (abs(m - k), m))[:views]
        t = time.perf_counter()
        ctx.set_views(2001 + k, poses[k], [2001 + m for m in nb], [poses[m] for m in nb])
        ctx.set_mask(2000)
        ctx.random_initialization(p)
        acc = ctx.preinit_warp(p) if k > 0 else 0
        ctx.red_black_pass(p, 0, iterations)
        ctx.finalize(p)
        ctx.memorize_result()
        ctx.synchronize()
        lat.append(1e3 * (time.perf_counter() - t))
        print(f"[profile_run] warm keyframe {k}: {lat[-1]:.2f} ms, {acc} warped candidates accepted", flush=True)
    if n > 1:
        print(f"[profile_run] warm start at {iterations} iterations, {views} views: p50 "
              f"{np.percentile(lat[1:], 50):.2f} ms, p99 {np.percentile(lat[1:], 99):.2f} ms", flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=960)
    ap.add_argument("--views", type=int, default=6)
    ap.add_argument("--pipeline", type=int, default=0, help="keyframes of the densify -> clean -> fuse trajectory")
    ap.add_argument("--batched", type=int, default=0, help="keyframes of the batched workload")
    ap.add_argument("--params", type=str, default="", help="JSON params config (params.params_from_json)")
    ap.add_argument("--warm", type=int, default=0, help="keyframes of the warm-started trajectory")
    ap.add_argument("--warm-iterations", type=int, default=1)
    args = ap.parse_args(argv)

    from patchmatch_amd import synth
    from patchmatch_amd.params import config1_params, params_from_json
    kf = synth.config1(0, W=args.width, H=args.height, n_views=args.views)
    print(f"[profile_run] rendered {args.width}x{args.height}+{args.views}", flush=True)

    from patchmatch_amd._lib import Context
    p = params_from_json(args.params) if args.params else config1_params()
    with Context(0) as ctx:
        ctx.load_keyframe(kf.ref, kf.views, kf.ref_pose, kf.view_poses)
        ctx.image_put(1000, kf.mask)
        ctx.set_mask(1000)
        ctx.set_timing(True)   # direct launches (no hipGraph): one dispatch per pass for the profiler
        for s in range(args.steps):
            t = time.perf_counter()
            ctx.densify(p)
            ctx.synchronize()
            e, c = ctx.eval_counts()
            print(f"[profile_run] step {s}: {1e3 * (time.perf_counter() - t):.1f} ms, {e} evaluations", flush=True)
        if args.batched > 0:
            from patchmatch_amd import workloads
            st = workloads.Stream(1, workloads.BATCH_CLASSES)
            r = workloads.run_batched(st, p, args.batched, 2)
            print(f"[profile_run] batched: {r['plain_depth_maps_per_s']} plain, {r['batched_depth_maps_per_s']} "
                  f"batched depth maps/s, config {r['autotuned_config']}", flush=True)
        if args.pipeline > 0:
            from patchmatch_amd import workloads
            r = workloads.run_pipeline(5, args.pipeline, args.width, args.height, params=p)
            print(f"[profile_run] pipeline of {args.pipeline} keyframes: {r['seconds']}", flush=True)
        if args.warm > 0:
            warm_sequence(ctx, p, args.warm, args.width, args.height, args.views, args.warm_iterations)
        if args.copy_mib > 0:
            ctx.debug_copy(args.copy_mib << 20, 1)
            print(f"[profile_run] copy of {args.copy_mib} MiB done", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
