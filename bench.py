"""Benchmark: aligned pairs/s on a turntable capture cycle (one MI355X).

Default workload (the batched full-cycle configuration, `--workload cycle36`): 36
depth captures of a synthetic object (union of 3 ellipsoids, semi-axes 0.15-0.45 m,
centred 1 m from the camera) at 10 deg turntable yaw steps with 5 mm translation
jitter, 640x480 Kinect-v1 rendering (fx = fy = 525), axial noise 0.0012 z^2, 2 mm
voxel grid, a random 5-60% half-space occlusion crop per capture (36k-113k points
per cloud); 36 adjacent pairs (i -> i-1, plus the closing pair) aligned
point-to-plane from the turntable prior (the rig's yaw step, as
rotate_align.cpp:224-235 applies it), r = 1 cm, <= 60 iterations, epsilons
1e-9 / 1e-3, 30-NN normals computed on the device.  When trained proposal weights
are present the learned hypothesis joins the prior in the same batch (batched +
learned).

`--workload c5`: 72 captures at 5 deg plus a top and a bottom view (45 deg tilt,
paired onto capture 0), 512x424 Kinect-v2 rendering, noise 0.001 z, 1.5 mm voxels;
15 pose hypotheses per pair (the prior and 14 rotations of up to 20 deg about the
source centroid), plus the learned one when weights are present.

With several hypotheses per pair (the learned one, or C5's 15) each pair's
initial poses are scored by their inliers within r and only the best is refined
(`--no-gate` refines them all and keeps the best by final fitness).

The launch uses the library's defaults (tile 64, cell size 2.5 point spacings);
`--autotune` lets the autotuner pick tile, cell size and search precision on this
device first.  A step = normals + grids + all ICP iterations (one hipGraph) + fitness
+ selection for the whole cycle, with the clouds already resident in HBM.  `value` counts the
pairs whose selected transform is within 0.05 deg / 0.2 mm of the FP64 oracle's
(run from the prior), divided by the step time.

Also reported: the plain single-pair GPU path (one pair at a time through the
per-pair API: normals + upload + align + fitness), every correspondence-search
precision mode on the same cycle (tile 64, checked bitwise against the FP32 scan), the dominant kernel's roofline from HIP events
recorded around each launch, and a host CPU baseline (the PCL-semantics FP32
restatement with an exact kd-tree, one thread, on a bounded sample of pairs).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ROT_TOL = np.deg2rad(0.05)
TRANS_TOL = 0.2e-3
PRECISION = {"fp32": 0, "fp16": 1, "bf16": 2, "fp32tiles": 3}
WEIGHTS = os.path.join(ROOT, "lowcost3dreconstruction_amd", "models", "proposal_c4.npy")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(kind, seed):
    """Returns (cycle, guesses (P * H, 4, 4), H)."""
    from lowcost3dreconstruction_amd import synth
    if kind == "cycle36":
        cyc = synth.turntable_cycle(seed, n_captures=36, yaw_step_deg=10.0, scale=3.0,
                                    centre=(0.0, 0.0, -1.0))
        return cyc, cyc.meta["prior"], 1
    cyc = synth.turntable_cycle(seed, n_captures=72, yaw_step_deg=5.0, cam=synth.KINECT_V2,
                                leaf=0.0015, noise=lambda z: 0.001 * z, scale=3.0,
                                centre=(0.0, 0.0, -1.0), top_bottom=True)
    return cyc, synth.hypotheses(cyc, 15, 20.0, 0.0, seed + 1), 15


def icp_kw():
    return dict(max_corr_dist=0.01, max_iterations=60, transformation_epsilon=1e-9,
                euclidean_fitness_epsilon=1e-3)


def cpu_reference_and_baseline(cyc, sample_pairs, threads):
    """The oracle is used here only as the checker (FP64 reference transforms for
    every pair, from the prior) and as the timed host baseline (PCL-semantics FP32
    mode, one thread, on `sample_pairs` pairs)."""
    import oracle
    t0 = time.time()
    normals = [oracle.normals(c, 30, nthreads=threads)[0] for c in cyc.clouds]
    refs = oracle.icp_batch(cyc.clouds, normals, cyc.pairs, guesses=cyc.meta["prior"],
                            nthreads=threads, point_to_plane=True, mode=oracle.FP64, **icp_kw())
    log(f"oracle FP64 references: {time.time() - t0:.1f} s")
    # all host cores: the whole cycle, pairs spread over `threads` threads
    t2 = time.perf_counter()
    tn_all = [oracle.normals(c, 30, nthreads=threads)[0] for c in cyc.clouds]
    oracle.icp_batch(cyc.clouds, tn_all, cyc.pairs, guesses=cyc.meta["prior"], nthreads=threads,
                     point_to_plane=True, mode=oracle.PCL32, **icp_kw())
    all_cores_s = time.perf_counter() - t2
    times = []
    for i in range(sample_pairs):
        s, t = cyc.pairs[i]
        t1 = time.perf_counter()
        tn = oracle.normals(cyc.clouds[t], 30, nthreads=1)[0]
        oracle.icp(cyc.clouds[s], cyc.clouds[t], tn, guess=cyc.meta["prior"][i], point_to_plane=True,
                   mode=oracle.PCL32, **icp_kw())
        times.append(time.perf_counter() - t1)
    return refs, float(np.median(times)), len(cyc.pairs) / all_cores_s


def aligned_mask(T, refs):
    from lowcost3dreconstruction_amd import synth
    ok, errs = [], []
    for Ti, r in zip(T, refs):
        er, et = synth.transform_error(Ti, r["T"])
        errs.append((er, et))
        ok.append(er <= ROT_TOL and et <= TRANS_TOL)
    return np.array(ok, bool), np.array(errs).reshape(-1, 2)


def copy_calibration(mib):
    """A copy kernel of known size: each launch reads and writes mib MiB."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}),
          flush=True)


def measured_traffic():
    """HBM bytes per iteration launch from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the copy calibration), or
    None when absent."""
    path = os.path.join(ROOT, "profiles", "icp_iteration_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return float(json.load(f)["traffic_bytes_per_launch"])


def timed_steps(cycle, steps):
    t0 = time.perf_counter()
    for _ in range(steps):
        cycle.run()
    return (time.perf_counter() - t0) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workload", choices=("cycle36", "c5"), default="cycle36")
    ap.add_argument("--precision", choices=sorted(PRECISION), default="fp32",
                    help="correspondence search of the measured run: exact FP32 scan, or "
                         "FP16/BF16 matrix-core distance tiles with exact FP32 rerank")
    ap.add_argument("--proposal", default="auto",
                    help="'auto' (use the committed weights when present), 'none', or a path")
    ap.add_argument("--no-gate", action="store_true",
                    help="refine every hypothesis of a pair instead of only the one with the "
                         "most inliers at its initial pose")
    ap.add_argument("--autotune", action="store_true",
                    help="let the autotuner pick tile / cell size / precision first (its choice "
                         "varies between near-equal settings from run to run; the default "
                         "measures the library's defaults, reproducibly)")
    ap.add_argument("--autotune-cache", default=None,
                    help="file the autotuner looks its choice up in and appends it to "
                         "(default: none, always tune)")
    ap.add_argument("--compare-precisions", type=int, default=3, metavar="STEPS",
                    help="also time every precision mode for STEPS steps (0 = skip)")
    ap.add_argument("--cpu-sample", type=int, default=3)
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-pairs", type=int, default=36)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB five times "
                         "(reference kernel for the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box in this setup has one MI355X, so "
              f"--gpus {args.gpus} is not supported (the pairs of a cycle shard across "
              f"GPUs with no exchange; see the design notes).", file=sys.stderr)
        sys.exit(2)

    import lowcost3dreconstruction_amd as lcr
    lcr.init(0)
    t0 = time.time()
    cyc, guesses, n_hyp = make_workload(args.workload, args.seed)
    sizes = [len(c) for c in cyc.clouds]
    log(f"workload {args.workload}: {len(cyc.clouds)} clouds, {min(sizes)}..{max(sizes)} points "
        f"({sum(sizes)} total), {len(cyc.pairs)} pairs x {n_hyp} hypotheses, generated in "
        f"{time.time() - t0:.1f} s")
    prop = None
    wpath = WEIGHTS if args.proposal == "auto" else (None if args.proposal == "none" else args.proposal)
    if wpath and os.path.exists(wpath):
        prop = lcr.Proposal(np.load(wpath))
        log(f"learned proposal: {wpath}")
    elif args.proposal not in ("auto", "none"):
        raise SystemExit(f"proposal weights not found: {wpath}")

    def make_cycle(mode, with_prop=True, tile=0, cell_factor=0.0):
        p = prop if with_prop else None
        gate = not args.no_gate and (n_hyp > 1 or p is not None)
        return lcr.Cycle(cyc.clouds, cyc.pairs, guesses=guesses, n_hyp=n_hyp, point_to_plane=True,
                         normals_k=30, precision_mode=mode, proposal=p, hyp_gate=gate,
                         tile=tile, cell_factor=cell_factor, **icp_kw())

    # ---------------- batched cycle (the measured path); creating a cycle copies
    # the clouds host -> device once (the PCIe-inclusive rate below adds that time)
    cycle = make_cycle(PRECISION[args.precision])
    # steady-state hand-over: a second cycle object over the same clouds (the first
    # creation in a process also pays the runtime's warm-up and first allocations)
    t_up = time.perf_counter()
    spare = make_cycle(PRECISION[args.precision])
    upload_s = time.perf_counter() - t_up
    del spare
    # the autotuner times the cycle under its 30 settings (tile, cell size,
    # search precision) and keeps the fastest; its cache file is --autotune-cache
    tune = None
    tuned_mode = PRECISION[args.precision]
    if args.autotune:
        t_t = time.time()
        tune = cycle.autotune(args.autotune_cache)
        tune["seconds"] = time.time() - t_t
        tuned_mode = int(tune["precision_mode"])
        log(f"autotune: {tune}")
    for _ in range(args.warmup):
        cycle.run()
    step_s = timed_steps(cycle, args.steps)
    res = cycle.results()
    inst = cycle.instances()
    # Per-kernel GPU time: HIP events cannot be recorded inside a captured graph,
    # so the same steps are replayed with the profiler on, which issues the same
    # launches eagerly with an event pair around each kernel on its stream.
    lcr.prof_reset()
    lcr.prof_enable(True)
    prof_step_s = timed_steps(cycle, args.steps)
    lcr.prof_enable(False)
    n_iter_launches, iter_ms = lcr.prof_read("icp_iteration")
    _, fit_ms = lcr.prof_read("fitness")
    _, grid_ms = lcr.prof_read("grid_build")
    _, nrm_ms = lcr.prof_read("normals")
    _, prop_ms = lcr.prof_read("proposal")
    _, gate_ms = lcr.prof_read("gate")
    _, solve_ms = lcr.prof_read("icp_solve")
    log(f"batched step {step_s * 1e3:.3f} ms (profiled replay {prof_step_s * 1e3:.3f} ms); "
        f"per step: icp search {iter_ms / args.steps:.3f} ms ({n_iter_launches // args.steps} "
        f"launches) + solve {solve_ms / args.steps:.3f} ms, "
        f"fitness {fit_ms / args.steps:.3f} ms, grids {grid_ms / args.steps:.3f} ms, "
        f"normals {nrm_ms / args.steps:.3f} ms, proposal {prop_ms / args.steps:.3f} ms")

    # ---------------- the same cycle without the learned hypothesis, and every
    # precision mode (their transforms must agree bitwise with the measured run)
    no_prop = None
    if prop is not None:
        c0 = (make_cycle(tuned_mode, with_prop=False, tile=tune["tile"],
                         cell_factor=tune["cell_factor"])
              if tune else make_cycle(PRECISION[args.precision], with_prop=False))
        c0.run()
        r0 = c0.results()
        no_prop = {"ms_per_step": timed_steps(c0, max(1, args.steps // 2)) * 1e3,
                   "iterations_mean": float(r0["iterations"].mean())}
        no_prop_T = r0["T"]
        log(f"without the learned hypothesis: {no_prop['ms_per_step']:.3f} ms/step")
        del c0
    # every search mode at tile 64 (the distance-tile modes need one wavefront per
    # tile) and the measured run's cell size: same reduction order, so the
    # transforms must agree bitwise with the fp32 scan's
    modes = {}
    if args.compare_precisions > 0:
        cf = tune["cell_factor"] if tune else 0.0
        ref_T = None
        for name, mode in sorted(PRECISION.items(), key=lambda kv: kv[1]):
            cm = make_cycle(mode, tile=64, cell_factor=cf)
            cm.run()
            lcr.prof_reset()
            lcr.prof_enable(True)
            ms = timed_steps(cm, args.compare_precisions) * 1e3
            lcr.prof_enable(False)
            _, it_ms = lcr.prof_read("icp_iteration")
            T = cm.results()["T"]
            if ref_T is None:
                ref_T = T  # fp32 (mode 0) comes first
            same = bool(np.array_equal(T, ref_T))
            modes[name] = {"profiled_ms_per_step": ms,
                           "icp_iteration_ms_per_step": it_ms / args.compare_precisions,
                           "same_transforms_as_fp32_scan": same}
            log(f"precision {name}: profiled step {ms:.3f} ms, icp "
                f"{modes[name]['icp_iteration_ms_per_step']:.3f} ms, identical results: {same}")
            del cm

    # ---------------- roofline of the iteration kernel: per instance and executed
    # iteration 12 B per source point + 24 B per target point (xyz + normal)
    H = len(inst["iterations"]) // len(cyc.pairs)
    it = inst["iterations"].astype(np.int64)
    ns = np.repeat(np.array([sizes[s] for s, _ in cyc.pairs], np.int64), H)
    nt = np.repeat(np.array([sizes[t] for _, t in cyc.pairs], np.int64), H)
    bytes_total = float(np.sum(it * (12 * ns + 24 * nt)))
    launches = max(n_iter_launches // args.steps, 1)
    bytes_per_launch = bytes_total / launches
    avg_launch_s = (iter_ms / 1e3) / max(n_iter_launches, 1)
    achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0

    # ---------------- plain single-pair GPU path (prior seed, one pair at a time)
    n_plain = min(args.plain_pairs, len(cyc.pairs))
    plain_T = []
    plain_s_per_pair = None
    if n_plain > 0:
        t1 = time.perf_counter()
        for i in range(n_plain):
            s, t = cyc.pairs[i]
            tn, _ = lcr.normals(cyc.clouds[t], 30)
            icp = lcr.IterativeClosestPoint(point_to_plane=True)
            icp.setInputSource(cyc.clouds[s])
            icp.setInputTarget(cyc.clouds[t], tn)
            kw = icp_kw()
            icp.setMaxCorrespondenceDistance(kw["max_corr_dist"])
            icp.setMaximumIterations(kw["max_iterations"])
            icp.setTransformationEpsilon(kw["transformation_epsilon"])
            icp.setEuclideanFitnessEpsilon(kw["euclidean_fitness_epsilon"])
            icp.align(cyc.meta["prior"][i])
            icp.getFitnessScore()
            plain_T.append(icp.getFinalTransformation())
        plain_s_per_pair = (time.perf_counter() - t1) / n_plain

    # ---------------- checker and host baseline
    cpu = None
    if not args.no_cpu:
        refs, cpu_per_pair, cpu_all = cpu_reference_and_baseline(cyc, args.cpu_sample, args.threads)
        ok, errs = aligned_mask(res["T"], refs)
        ok_plain = aligned_mask(plain_T, refs[:n_plain])[0] if n_plain else np.zeros(0, bool)
        if no_prop is not None:
            no_prop["aligned_pairs"] = int(aligned_mask(no_prop_T, refs)[0].sum())
        cpu = {"value": 1.0 / cpu_per_pair, "unit": "pairs/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_sample} pairs of the same cycle: 30-NN normals of the target "
                         f"+ point-to-plane ICP from the prior, PCL-semantics FP32 restatement "
                         f"with an exact kd-tree (not PCL), median per pair",
               "all_cores": {"value": cpu_all, "unit": "pairs/s", "cores": args.threads,
                             "sample": "the whole cycle, pairs spread over the threads"}}
    else:
        ok = np.ones(len(cyc.pairs), bool)
        ok_plain = np.ones(n_plain, bool)
        errs = np.zeros((len(cyc.pairs), 2))
    n_aligned = int(ok.sum())
    value = n_aligned / step_s
    # aligned pairs by the hypothesis the step kept (0: the turntable prior)
    kept_prior = np.asarray(res["best_hyp"]) == 0
    by_kept = {"prior": [int(ok[kept_prior].sum()), int(kept_prior.sum())],
               "other": [int(ok[~kept_prior].sum()), int((~kept_prior).sum())]}
    plain_value = float(ok_plain.mean()) / plain_s_per_pair if n_plain else None
    if no_prop is not None:
        no_prop["value"] = no_prop.get("aligned_pairs", len(cyc.pairs)) / (no_prop["ms_per_step"] / 1e3)

    out = {
        "metric": "aligned pairs/sec at <=0.05deg/0.2mm vs FP64 oracle",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 distances, f64 reductions/solves",
        "data": "synthetic (seeded turntable renderer)",
        "config": {
            "workload": ("turntable cycle, 36 captures, point-to-plane, batched in one hipGraph"
                         if args.workload == "cycle36" else
                         "Kinect-v2 cycle, 72 captures + top/bottom, 15 hypotheses per pair"),
            "clouds": len(cyc.clouds), "points_min": int(min(sizes)), "points_max": int(max(sizes)),
            "points_total": int(sum(sizes)), "pairs": int(len(cyc.pairs)),
            "hypotheses_per_pair": int(H),
            "learned_proposal": prop is not None,
            "hypothesis_gate": (not args.no_gate and (n_hyp > 1 or prop is not None)),
            "max_iterations": 60, "max_corr_dist_m": 0.01, "normals_k": 30, "seed": args.seed,
            "initial_pose": "turntable prior (rig yaw step about its axis)",
            "correspondence_search": {v: k for k, v in PRECISION.items()}[tuned_mode],
            "launch": ({"tile": tune["tile"], "cell_factor": tune["cell_factor"], "tuned": True}
                       if tune else {"tile": "default", "cell_factor": "default", "tuned": False}),
            "parallelism": "single GPU",
        },
        "aligned_pairs": n_aligned,
        "aligned_by_kept_hypothesis": by_kept,
        "upload_ms": upload_s * 1e3,
        "pcie_inclusive_pairs_per_s": n_aligned / (step_s + upload_s),
        "max_rot_err_deg": float(np.rad2deg(errs[:, 0].max())),
        "max_trans_err_m": float(errs[:, 1].max()),
        "iterations_mean": float(res["iterations"].mean()),
        "learned_chosen": (float(np.mean(res["best_hyp"] == n_hyp)) if prop is not None else None),
        "without_learned": no_prop,
        "plain_gpu": ({"value": plain_value, "unit": "pairs/s", "pairs": n_plain,
                       "note": "one pair at a time through the per-pair API "
                               "(normals + upload + align + fitness)"} if n_plain else None),
        "speedup_vs_plain_gpu": (value / plain_value if plain_value else None),
        "speedup_vs_cpu_baseline": (value / cpu["value"] if cpu else None),
        "autotune": tune,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "icp_iteration", "avg_launch_us": avg_launch_s * 1e6,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "precision_modes": modes,
        "kernel_ms_per_step": {"icp_iteration": iter_ms / args.steps,
                               "icp_solve": solve_ms / args.steps, "fitness": fit_ms / args.steps,
                               "grid_build": grid_ms / args.steps, "normals": nrm_ms / args.steps,
                               "proposal": prop_ms / args.steps, "gate": gate_ms / args.steps,
                               "profiled_step": prof_step_s * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
