"""Benchmark: keyframe densification on one MI355X (plain drop-in form).

Workload (the configuration the headline metric is quoted on): one
1920x960 uint8 equirectangular reference keyframe + 6 neighbours (baselines
0.05-0.5 m) rendered from a seeded synthetic scene (box room, 20 spheres,
1.2 m tube, 4-octave value-noise textures of 4 to 32 cells per metre,
photometric jitter, bottom 15 % masked);
11x11 patch at stride 2 (36 taps), best-3-of-6 mean NCC, 8 red-black
iterations, FP32.  A step is one full densification of that keyframe
(random init + 16 red-black passes + copy-back + PostProcess) with images and
mask resident in HBM.

Prints one JSON line (rank 0).  Extra fields: latency percentiles, accuracy
against the scene's ground truth, the dominant kernel's roofline position and
the CPU oracle timed on a bounded sample on this host; then the two other
forms of the metric (patchmatch_amd/workloads.py):
  "batched": 64 keyframes of mixed resolution (960x480 / 1920x960 /
             2880x1440, weights 0.3/0.5/0.2) and view counts U{2..10}, batched
             (one work-queue launch per stage, hipGraph, autotuned) vs plain;
  "served":  Poisson request streams at 5, 15 and 30 requests/s drawn from
             a 500-keyframe trajectory (125 keyframes per resolution class,
             up to 5760x2880, views U{2..10}), served (continuous batching,
             LRU image pool) vs the plain per-request path, p50/p99 latency
             from arrival to maps on the host; each form's capacity (all
             requests at once) and a rate at half the served capacity; every
             request of a stream checked bit for bit against plain under an
             HBM budget that evicts images;
  "served_1920x960": the same at 1920x960 with 6 neighbours only;
  "pipeline": 30 keyframes of one trajectory at the plain size, densified
             against their 4-8 nearest neighbours, cleaned (5-map window,
             depth within 1 %, round trip within 1 px, normals within 10
             degrees) and fused (Pruner): accuracy after each stage; then
             densified once more one by one with the densify thread's warm
             start (the previous keyframe's result warped in), against the
             plain workload's neighbour count, at 8, 4, 2 and 1 iterations,
             with each keyframe's latency.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3            # MI355X published FP32 vector peak (counts packed FP32 FMA)
FP32_NONPACKED_PEAK_TFLOPS = 78.6   # the same without packed FP32, which this build does not use
HBM_PEAK_GBS = 8000.0
# L2-to-memory bytes per k_propagate_q launch on the 1920x960 + 6-view
# workload, from rocprofv3 FETCH_SIZE and WRITE_SIZE scaled by a copy of known
# size (profiles/k_propagate_q_c1_1920x960_6views.md); None for other workloads.
TRAFFIC_BYTES_PER_LAUNCH_C1 = int((181_613 * 2.00 + 43_038 * 1.00) * 1024)
TRAFFIC_SOURCE = ("profile-derived, not measured in this run: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of "
                  "k_propagate_q on this workload, scaled by a streaming copy of known size "
                  "(profiles/k_propagate_q_c1_1920x960_6views.md)")
# VALU wave-instructions per k_propagate_q launch on the 1920x960 + 6-view
# workload (rocprofv3 --pmc SQ_INSTS_VALU, same profile): the issue side of
# the roofline, against one wave64 VALU instruction per CU per clock.
VALU_WAVE_INSTS_PER_LAUNCH_C1 = 2.539e9
CUS, CLOCK_HZ = 256, 2.4e9


def ops_per_eval(T: int, N: int) -> int:
    """Algorithmic FP32 operations of one plane-hypothesis score: T taps x
    (19 view-independent + 52 per view) + 25 per view finalise + 10."""
    return T * (19 + 52 * N) + 25 * N + 10


def cpu_baseline(kf, params, init_maps, target_s: float, threads: int):
    """Oracle (host C/OpenMP restatement) on a bounded sample of the same
    keyframe: whole red-black passes, in order from the same initial maps,
    until target_s seconds have elapsed (at least one pass).  Returns
    hypothesis evaluations per second and a description of the sample."""
    from oracle import or_py
    or_py.set_threads(threads)
    ok = or_py.Keyframe.from_synth(kf)
    plane, depth, score = (np.array(a, copy=True) for a in init_maps)
    evals, elapsed, passes = 0, 0.0, 0
    while passes < 2 * params.iterations and (passes == 0 or elapsed < target_s):
        it, colour = divmod(passes, 2)
        t = time.perf_counter()
        evals += or_py.propagate_rows(ok, params, plane, depth, score, it, colour, 0, kf.H)
        elapsed += time.perf_counter() - t
        passes += 1
    return evals / elapsed, dict(passes=passes, evals=evals, seconds=elapsed)


def cpu_c0_end_to_end(threads: int, repeats: int = 5):
    """Oracle densification of the 480x240 + 2-view box-room keyframe (7x7
    patch, 4 iterations) end to end: median seconds of `repeats` runs."""
    from oracle import or_py
    from patchmatch_amd import synth
    from patchmatch_amd.params import config0_params
    or_py.set_threads(threads)
    kf = synth.config0(0)
    ok = or_py.Keyframe.from_synth(kf)
    p = config0_params()
    ts = []
    for _ in range(repeats):
        t = time.perf_counter()
        or_py.densify(ok, p, kf.mask)
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def gpu_c0_end_to_end(ctx, repeats: int = 5):
    """The same c0 densification on the GPU (images already resident):
    median milliseconds of `repeats` runs after one warm-up."""
    from patchmatch_amd import synth
    from patchmatch_amd.params import config0_params
    kf = synth.config0(0)
    p = config0_params()
    ctx.load_keyframe(kf.ref, kf.views, kf.ref_pose, kf.view_poses, base_id=2000)
    ctx.image_put(2999, kf.mask)
    ctx.set_mask(2999)
    ctx.densify(p)
    ctx.synchronize()
    ts = []
    for _ in range(repeats):
        t = time.perf_counter()
        ctx.densify(p)
        ctx.synchronize()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)) * 1e3


def reference_defaults_keyframe():
    """The reference's default operating point (example/patchmatch/config.yaml:39-44,
    depthmap_computation_module.cc:211): a 640x320 depth map against 2
    neighbours, 36-tap patch; box-room scene as in config 0."""
    from patchmatch_amd import synth
    from patchmatch_amd.params import params_from_json
    kf = synth.make_keyframe(7, 640, 320, 2, baseline=(0.1, 0.3))
    return kf, params_from_json({"base": "default", "iterations": 4})


def reference_defaults_times(ctx, threads: int, repeats: int = 5):
    """Median end-to-end densification time of that keyframe with pmp_default_params
    (the reference's constants, pooled view fold) and 4 iterations, on the GPU
    (images resident) and in the oracle on `threads` host threads."""
    from oracle import or_py
    kf, p = reference_defaults_keyframe()
    ctx.load_keyframe(kf.ref, kf.views, kf.ref_pose, kf.view_poses, base_id=3000)
    ctx.image_put(3999, kf.mask)
    ctx.set_mask(3999)
    ctx.densify(p)
    ctx.synchronize()
    g = []
    for _ in range(repeats):
        t = time.perf_counter()
        ctx.densify(p)
        ctx.synchronize()
        g.append(time.perf_counter() - t)
    or_py.set_threads(threads)
    ok = or_py.Keyframe.from_synth(kf)
    c = []
    for _ in range(3):
        t = time.perf_counter()
        or_py.densify(ok, p, kf.mask)
        c.append(time.perf_counter() - t)
    return {"workload": "640x320 + 2 neighbours, 11x11 patch at stride 2, pooled view fold, 4 iterations "
                        "(the reference's example configuration), box-room scene",
            "gpu_ms": round(float(np.median(g)) * 1e3, 3), "oracle_s": round(float(np.median(c)), 3),
            "oracle_threads": threads}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=960)
    ap.add_argument("--views", type=int, default=6)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample length (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, available cores)")
    ap.add_argument("--batched", type=int, default=64, help="keyframes of the batched workload (0: skip)")
    ap.add_argument("--served-requests", type=int, default=24, help="requests per served rate (0: skip)")
    ap.add_argument("--rates", type=str, default="5,15,30", help="served arrival rates (requests/s)")
    ap.add_argument("--pipeline", type=int, default=30,
                    help="keyframes of the densify -> clean -> fuse trajectory at the plain size (0: skip)")
    ap.add_argument("--params", type=str, default="",
                    help="JSON params config for every leg (params.params_from_json; default: config 1 with --seed)")
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build measures one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not implemented)"}))
        return 2

    from patchmatch_amd import synth
    from patchmatch_amd.params import config1_params, params_from_json

    t0 = time.perf_counter()
    kf = synth.config1(args.seed, W=args.width, H=args.height, n_views=args.views)   # before any GPU use
    t_render = time.perf_counter() - t0
    print(f"[bench] rendered {args.width}x{args.height} + {args.views} views in {t_render:.1f}s", file=sys.stderr,
          flush=True)

    from patchmatch_amd._lib import PMP_MAPS_FINAL, PMP_MAPS_INIT, Context
    p = params_from_json(args.params) if args.params else config1_params(seed=args.seed)
    ctx = Context(0)
    ctx.load_keyframe(kf.ref, kf.views, kf.ref_pose, kf.view_poses)
    mask_id = 1000
    ctx.image_put(mask_id, kf.mask)
    ctx.set_mask(mask_id)

    for _ in range(args.warmup):
        ctx.densify(p)
    ctx.synchronize()
    print(f"[bench] warmup {args.warmup} done", file=sys.stderr, flush=True)

    ctx.set_timing(True)
    lat = []
    evals = cands = 0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        ctx.densify(p)
        ctx.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
        e, c = ctx.eval_counts()
        evals += e
        cands += c
    ctx.synchronize()
    wall = time.perf_counter() - t_start
    tm = ctx.get_timing()
    ctx.set_timing(False)
    print(f"[bench] timed {args.steps} steps in {wall:.2f}s", file=sys.stderr, flush=True)

    # accuracy of the last result against ground truth (unmasked, valid pixels)
    _, depth, score, _ = ctx.get_maps(PMP_MAPS_FINAL)
    valid = (kf.mask > 0) & (depth > 0)
    unmasked = kf.mask > 0
    rel = np.abs(depth.astype(np.float64) - kf.gt_depth) / kf.gt_depth
    acc = dict(valid_fraction=round(float(valid.sum() / unmasked.sum()), 4),
               absrel_le_1pct_of_valid=round(float(np.mean(rel[valid] <= 0.01)), 4),
               absrel_le_1pct_of_unmasked=round(float(np.sum(valid & (rel <= 0.01)) / unmasked.sum()), 4),
               median_absrel=round(float(np.median(rel[valid])), 5))

    T = len(range(-p.patch_half, p.patch_half + 1, p.patch_stride)) ** 2
    N = args.views
    P_int = (args.height - 2 * p.patch_half - 1) * (args.width - 2 * p.patch_half - 1)
    launches = tm["propagate_launches"]
    avg_launch_ms = tm["propagate_ms"] / max(launches, 1)
    evals_per_launch = evals / max(launches, 1)
    opc = ops_per_eval(T, N)
    achieved_tflops = evals_per_launch * opc / (avg_launch_ms * 1e-3) / 1e12
    # the reference scores all 8 candidates and every refinement of every
    # pixel of the pass's colour (no pre-rejection): half the interior pixels
    ref_evals_per_launch = P_int * (8 + p.n_refine) / 2
    ref_equiv_tflops = ref_evals_per_launch * opc / (avg_launch_ms * 1e-3) / 1e12
    ms_per_step = wall * 1e3 / args.steps
    value = args.steps / wall
    mpx_iter = P_int * p.iterations / 1e6

    result = {
        "metric": "depth maps/s + p50/p99 keyframe latency, served vs plain, at fixed accuracy",
        "value": round(value, 4),
        "unit": "depth maps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded ray-cast scene, rendered in-process)",
        "config": {"workload": "plain drop-in: 1 keyframe 1920x960 + 6 neighbours, 11x11 patch stride 2, "
                               "best-3-of-6 NCC, 8 red-black iterations"
                   if (args.width, args.height, N) == (1920, 960, 6) and not args.params
                   else f"plain drop-in: 1 keyframe {args.width}x{args.height} + {N} neighbours"
                        + (f", params from {args.params}" if args.params else ""),
                   "width": args.width, "height": args.height, "views": N, "taps": T, "iterations": p.iterations,
                   "texture_cells_per_m": [synth.TEX_FREQ_C1, 8 * synth.TEX_FREQ_C1],
                   "form": "plain", **({"params": p.as_dict()} if args.params else {})},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                       "max": round(float(np.max(lat)), 3)},
        "mpx_iter_per_s": round(mpx_iter * value, 2),
        "stage_ms_per_step": {"init": round(tm["init_ms"] / args.steps, 3),
                              "propagate": round(tm["propagate_ms"] / args.steps, 3),
                              "finalize": round(tm["finalize_ms"] / args.steps, 3)},
        "evals_per_step": int(evals / args.steps),
        "candidates_per_step": int(cands / args.steps),
        "accuracy_vs_ground_truth": acc,
        "roofline": {"bound": "valu", "kernel": "k_propagate_q", "achieved": round(achieved_tflops, 3),
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved_tflops / FP32_PEAK_TFLOPS, 4),
                     "peak_nonpacked": FP32_NONPACKED_PEAK_TFLOPS,
                     "frac_nonpacked": round(achieved_tflops / FP32_NONPACKED_PEAK_TFLOPS, 4),
                     "traffic": TRAFFIC_BYTES_PER_LAUNCH_C1 if (args.width, args.height, N) == (1920, 960, 6) else None,
                     "traffic_unit": "bytes per launch (L2 misses)", "traffic_source": TRAFFIC_SOURCE,
                     "ops_model": "T (19 + 52 N) + 25 N + 10 FP32 ops per hypothesis evaluation "
                                  "(FMA 2, add/mul/compare 1, each elementary function 1); executed evaluations "
                                  "only (pre-rejected candidates are not scored)",
                     "ops_per_eval": opc, "evals_per_launch": int(evals_per_launch),
                     "avg_launch_ms": round(avg_launch_ms, 4), "launches": launches,
                     "valu_issue": ({
                         "source": "SQ_INSTS_VALU per launch from the same profile (not measured in this run), "
                                   "over this run's launch time",
                         "lane_instructions_per_sample": round(VALU_WAVE_INSTS_PER_LAUNCH_C1 * 64 /
                                                               (evals_per_launch * T * N), 1),
                         "wave_instructions_per_s": round(VALU_WAVE_INSTS_PER_LAUNCH_C1 / (avg_launch_ms * 1e-3), -8),
                         "peak_wave_instructions_per_s": CUS * CLOCK_HZ,
                         "peak_note": "one wave64 VALU instruction per CU per clock at 2.4 GHz (the non-packed "
                                      "FP32 FMA rate); the profile's issue rate per SIMD is the measured ceiling",
                         "frac": round(VALU_WAVE_INSTS_PER_LAUNCH_C1 / (avg_launch_ms * 1e-3) / (CUS * CLOCK_HZ), 3)}
                         if (args.width, args.height, N) == (1920, 960, 6) else None),
                     "reference_equivalent": {
                         "note": "every candidate and refinement of every pixel scored, as the reference does "
                                 "(P (8 + n_refine) / 2 evaluations per pass); not what this kernel executes",
                         "evals_per_launch": int(ref_evals_per_launch),
                         "achieved": round(ref_equiv_tflops, 3),
                         "frac": round(ref_equiv_tflops / FP32_PEAK_TFLOPS, 4),
                         "frac_nonpacked": round(ref_equiv_tflops / FP32_NONPACKED_PEAK_TFLOPS, 4)}},
    }

    if args.cpu_seconds > 0:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        # OpenMP reads these when the oracle (and with it libgomp) is first
        # loaded, which is here: threads bound to cores, close together
        os.environ.setdefault("OMP_PROC_BIND", "close")
        os.environ.setdefault("OMP_PLACES", "cores")
        init_maps = ctx.get_maps(PMP_MAPS_INIT)
        eps, sample = cpu_baseline(kf, p, init_maps, args.cpu_seconds, threads)
        init_evals = P_int
        evals_per_keyframe = evals / args.steps + init_evals
        cpu_value = eps / evals_per_keyframe
        result["cpu_baseline"] = {
            "value": round(cpu_value, 6), "unit": "depth maps/s", "cores": threads, "kind": "port",
            "sample": (f"oracle: the first {sample['passes']} of {2 * p.iterations} red-black passes of the "
                       f"same keyframe from the same initial maps ({sample['evals']} hypothesis evaluations in "
                       f"{sample['seconds']:.1f} s); depth maps/s = evaluations/s / {int(evals_per_keyframe)} "
                       f"evaluations per keyframe (init + passes, as counted on the GPU)"),
            "evals_per_s": round(eps, 1),
            "tne_per_s": round(eps * T * N, 1),
            "mpx_iter_per_s": round(mpx_iter * cpu_value, 4),
            "omp": {"proc_bind": os.environ["OMP_PROC_BIND"], "places": os.environ["OMP_PLACES"]},
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "c0_end_to_end_s": round(cpu_c0_end_to_end(threads), 3),
            "c0_gpu_end_to_end_ms": round(gpu_c0_end_to_end(ctx), 3),
            "c0_note": "480x240 + 2 views, 7x7 patch, 4 iterations: median of 5 densifications, oracle and GPU",
            "reference_defaults": reference_defaults_times(ctx, threads)}
    ctx.close()

    if args.batched > 0 or args.served_requests > 0 or args.pipeline > 0:
        from patchmatch_amd import workloads
        if args.pipeline > 0:
            t = time.perf_counter()
            result["pipeline"] = workloads.run_pipeline(args.seed + 5, args.pipeline, args.width, args.height,
                                                        params=p, warm_iterations=(4, 2, 1), warm_views=args.views)
            print(f"[bench] pipeline in {time.perf_counter() - t:.1f}s: {result['pipeline']}", file=sys.stderr,
                  flush=True)
        if args.batched > 0:
            t = time.perf_counter()
            st = workloads.Stream(args.seed + 1, workloads.BATCH_CLASSES)
            print(f"[bench] batched workload rendered in {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
            result["batched"] = workloads.run_batched(st, p, args.batched, args.seed + 2)
            print(f"[bench] batched: {result['batched']}", file=sys.stderr, flush=True)
            del st
        if args.served_requests > 0:
            t = time.perf_counter()
            # 500 keyframes: 125 along the seeded trajectory per resolution class
            st = workloads.Stream(args.seed + 3, workloads.SERVE_CLASSES, per_class=125, with_gt=False)
            print(f"[bench] served workload rendered in {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
            rates = [float(r) for r in args.rates.split(",") if r]
            result["served"] = workloads.run_served(st, p, rates, args.served_requests, args.seed + 4, tail_stages=2)
            print(f"[bench] served: {result['served']}", file=sys.stderr, flush=True)
            # every request of a stream against plain, under a budget that holds the maps of two
            # 5760x2880 slots and the images of two requests. Requests 0 and 1 are 5760x2880 so that
            # both slots grow at once; with 8 and 13 they cover four disjoint 4-neighbour image sets,
            # more than that room, so pooled images must be evicted.
            specs = st.draw(np.random.default_rng(args.seed + 6), 16)
            c_big = len(st.classes) - 1
            for q, k in ((0, 10), (1, 40), (8, 70), (13, 100)):
                specs[q] = (c_big, k, [k - 2, k - 1, k + 1, k + 2])
            big = st.classes[c_big][0] * st.classes[c_big][1]
            total, one = workloads.image_bytes(st, specs)
            room = 2 * one + (64 << 20)
            result["served"]["verification"] = workloads.verify_served(
                st, p, specs, max_slots=2, hbm_budget=2 * 84 * big + room, window=2)
            result["served"]["verification"].update(image_bytes=total, image_room=room)
            print(f"[bench] served verification: {result['served']['verification']}", file=sys.stderr, flush=True)
            del st
            # the plain workload's shape alone (1920x960, 6 neighbours) below capacity: the latency target's case
            st = workloads.Stream(args.seed + 7, [(1920, 960, 1.0)], per_class=16, with_gt=False)
            result["served_1920x960"] = workloads.run_served(st, p, [], args.served_requests, args.seed + 8,
                                                             views=(6, 6), tail_stages=2)
            print(f"[bench] served 1920x960: {result['served_1920x960']}", file=sys.stderr, flush=True)
            del st
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
