"""Benchmark: plain GPU PatchMatch Mpix/s on the single-view workload.

Workload (one "step" = one full ComputePatchMatch of the reference view):
1600x1200 uint8 reference view + 8 neighbours rendered from the seeded
scene generator (20 boxes on a sinusoidal terrain of amplitude 1.5 and
period 6), depth range [1, 40], patch 7x7, min patch SD 1, FP32 samples and
FP32 NCC.  Inputs are resident in HBM when the timed region starts (views
uploaded by add_view); each step runs the schedule that reaches the
reference's 5-iteration accuracy and mean score (checked over seeds 1-5
against the oracle's recorded figures): a 4-level pyramid (the views
box-averaged by 8, 4 and 2 with 2, 2 and 1 iterations), each level's planes
seeding the next, then 2 full-resolution iterations with the
median-consistent re-selection and the median post-process, into
device-resident output maps.
Plain PatchMatch at the reference's 5 iterations is timed beside it.

Prints one JSON line.  `roofline` is for the dominant kernel (the red/black
PatchMatch half-iteration), timed with HIP events on the stream it runs on;
its algorithmic work is 28 flop per bilinear NCC sample plus 160 flop of
setup per scoring, over the scorings the kernel actually executed (a device
counter).  `cpu_baseline` times the oracle (the C restatement of the
reference's CPU ComputePatchMatch) on bounded crops of the same scene, one
crop per host thread, before the GPU is touched.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

FP32_PEAK_TFLOPS = 157.3   # MI355X dense FP32 vector (published)
FLOP_PER_SAMPLE = 28
FLOP_PER_SCORING = 160

WORKLOAD = dict(width=1600, height=1200, neighbours=8, boxes=20,
                terrain=(1.5, 6.0), min_depth=1.0, max_depth=40.0, patch=7,
                min_patch_sd=1.0, iterations=5, seed=2000,
                pyramid=(2, 2, 1), full_iterations=2, coarse_reselection=False)


GOLDEN_CFG2 = os.path.join(HERE, "tests", "golden", "config2_oracle_fullview.json")
PMC_SUMMARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                           "pm_half_p7_1600x1200_pmc.md")


def pmc_traffic():
    """HBM bytes per k_pm_half<7> launch on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a copy
    kernel of known size); None if the summary is absent."""
    try:
        for line in open(PMC_SUMMARY):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_scene():
    from opensfm_amd import scenes
    w = WORKLOAD
    return scenes.scene_config2(seed=w["seed"], width=w["width"], height=w["height"])


def crop_views(scene, x0, y0, cw, ch):
    """Reference view cropped to (x0, y0, cw, ch) with K shifted; neighbours
    unchanged, so the per-pixel work equals the full view's."""
    views = []
    for k, v in enumerate(scene.views):
        if k == 0:
            K = v.K.copy()
            K[0, 2] -= x0
            K[1, 2] -= y0
            img = np.ascontiguousarray(v.image[y0:y0 + ch, x0:x0 + cw])
            views.append((K, v.R, v.t, img, np.full_like(img, 255)))
        else:
            views.append((v.K, v.R, v.t, v.image, v.mask))
    return views


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, threads, seed_sets=5, target_s=4.0):
    """Oracle ComputePatchMatch on crops of the reference view, one crop per
    thread (ctypes releases the GIL), for `seed_sets` sets of RNG seeds; each
    set takes about target_s seconds."""
    import oracle
    w = WORKLOAD

    def run(x0, y0, cw, ch, seed, keep=None):
        est = oracle.OracleEstimator()
        for K, R, t, img, mask in crop_views(scene, x0, y0, cw, ch):
            est.add_view(K, R, t, img, mask)
        est.set_depth_range(w["min_depth"], w["max_depth"], 50)
        est.set_patch_size(w["patch"])
        est.set_min_patch_sd(w["min_patch_sd"])
        est.set_patchmatch_iterations(w["iterations"])
        est.set_seed(seed)
        t0 = time.perf_counter()
        depth = est.compute_patch_match()[0]
        dt = time.perf_counter() - t0
        if keep is not None:
            keep.append(depth)
        return dt

    # calibrate on a small crop, then size one crop per thread for target_s
    t_small = run(700, 500, 40, 40, 1)
    per_px = t_small / ((40 - 6) * (40 - 6))
    side = int(np.clip(np.sqrt(target_s / per_px), 40, 400))
    rng = np.random.default_rng(0)
    jobs = [(int(rng.integers(0, w["width"] - side)), int(rng.integers(0, w["height"] - side)))
            for _ in range(threads)]
    rates, one_thread, crops = [], [], []
    for ss in range(seed_sets):
        res = [None] * threads
        depths = [[] for _ in range(threads)]

        def worker(k):
            res[k] = run(jobs[k][0], jobs[k][1], side, side, 100 + 1000 * ss + k, depths[k])

        t0 = time.perf_counter()
        ths = [threading.Thread(target=worker, args=(k,)) for k in range(threads)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        wall = time.perf_counter() - t0
        rates.append(threads * side * side / wall / 1e6)
        one_thread.append(side * side / np.mean(res) / 1e6)
        if ss == 0:
            crops = [(x0, y0, side, side, depths[k][0]) for k, (x0, y0) in enumerate(jobs)]
    return crops, {
        "value": float(np.mean(rates)), "unit": "Mpix/s", "cores": threads, "kind": "port",
        "sample": (f"{threads} concurrent {side}x{side} crops of the 1600x1200 reference view "
                   f"(8 neighbours, 5 iterations, patch 7), one per host thread, {seed_sets} "
                   f"seed sets (Mpix/s {min(rates):.4f}-{max(rates):.4f}); 1-thread rate "
                   f"{np.mean(one_thread):.5f} Mpix/s; oracle built gcc -O3 -ffp-contract=off"),
        "cpu_model": cpu_model(), "host_cpus_visible": len(os.sched_getaffinity(0)),
        "wall_s": float(sum(threads * side * side / r / 1e6 for r in rates)),
    }


def street_crops_oracle(sc, patch, iterations, gpu_schedule, jobs=16, side=160, seed=4000,
                        seed_sets=2):
    """Accuracy reference for the street batch: the oracle's ComputePatchMatch
    at `iterations` and the GPU engine with gpu_schedule = (iterations per
    coarser level, coarsest first; full iterations; re-selection at the
    coarser levels) on the same crops (side x side from the
    centre of the first `jobs` reference views, K shifted, all of the job's
    neighbours), for `seed_sets` sets of seeds, one oracle run per host
    thread."""
    import oracle
    from opensfm_amd import DepthmapEstimator, scenes
    crops = []
    for k, (ref, nb) in enumerate(sc.jobs[:jobs]):
        R = sc.views[ref]
        H, W = R.image.shape
        x0, y0 = (W - side) // 2, (H - side) // 2
        K = R.K.copy()
        K[0, 2] -= x0
        K[1, 2] -= y0
        img = np.ascontiguousarray(R.image[y0:y0 + side, x0:x0 + side])
        views = [(K, R.R, R.t, img, np.full_like(img, 255))]
        views += [(sc.views[v].K, sc.views[v].R, sc.views[v].t, sc.views[v].image,
                   sc.views[v].mask) for v in nb]
        crops.append((views, R.depth[y0:y0 + side, x0:x0 + side]))

    def setup(est, views, s):
        for v in views:
            est.add_view(*v)
        est.set_depth_range(sc.min_depth, sc.max_depth, 50)
        est.set_patch_size(patch)
        est.set_min_patch_sd(1.0)
        est.set_patchmatch_iterations(iterations)
        est.set_seed(s)
        return est

    runs = [(k, seed + 1000 * ss + k) for ss in range(seed_sets) for k in range(len(crops))]
    cpu = [None] * len(runs)
    nthreads = max(1, min(16, len(os.sched_getaffinity(0))))

    def acc(d, k):
        """(whole crop, at least 20 px from its edges)"""
        gt = crops[k][1]
        return (scenes.depth_accuracy(d, gt, sc.min_depth, sc.max_depth),
                scenes.depth_accuracy(d[20:-20, 20:-20], gt[20:-20, 20:-20], sc.min_depth,
                                      sc.max_depth))

    def worker(w):
        for r in range(w, len(runs), nthreads):
            k, s = runs[r]
            cpu[r] = acc(setup(oracle.OracleEstimator(), crops[k][0], s).compute_patch_match()[0], k)

    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker, args=(w,)) for w in range(nthreads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    cpu_s = time.perf_counter() - t0
    def gpu_est(views, s):
        e = setup(DepthmapEstimator(), views, s)
        e.set_pyramid(gpu_schedule[0])
        e.set_patchmatch_iterations(gpu_schedule[1])
        e.set_pyramid_consistency(gpu_schedule[2])
        return e

    gpu = [acc(gpu_est(crops[k][0], s).compute_patch_match()[0], k) for k, s in runs]
    cpu, gpu = np.array(cpu), np.array(gpu)
    diff = gpu - cpu
    se = diff.std(axis=0, ddof=1) / np.sqrt(len(diff))
    return {"crops": len(crops), "side": side, "seed_sets": seed_sets,
            "oracle_accuracy": round(float(cpu[:, 0].mean()), 4),
            "gpu_accuracy": round(float(gpu[:, 0].mean()), 4),
            "gpu_minus_oracle_se": round(float(se[0]), 4),
            "oracle_accuracy_interior": round(float(cpu[:, 1].mean()), 4),
            "gpu_accuracy_interior": round(float(gpu[:, 1].mean()), 4),
            "gpu_minus_oracle_se_interior": round(float(se[1]), 4),
            "gpu_at_least_oracle": bool(gpu[:, 0].mean() >= cpu[:, 0].mean()),
            "oracle_s": round(cpu_s, 1)}


# street batch: iterations per coarser level (1/8, 1/4, 1/2), full iterations,
# the median-consistent re-selection at the coarser levels
BATCH_SCHEDULE = ((2, 2, 1), 2, False)


def batch_bench(n_views, steps=3, seed=4000, cpu_baseline=True):
    """Batched street workload: n_views reference views of sizes drawn from
    {640x480, 1024x768, 1600x1200, 2048x1536}, 4-12 neighbours each
    (rendered on the GPU), a 4-level pyramid with 2, 2, 1 and 2 iterations
    (coarsest first; the reference's count is 3; accuracy checked against the
    oracle's 3 iterations on crops) in one captured hipGraph, tile
    autotuned by speed and patch size {5, 7} gated on accuracy (the fastest
    within 0.2 points of the most accurate), then the cleaner with 2
    consistent views.  Plain 3-iteration PatchMatch is timed beside it.  Compared with one DepthmapEstimator run per view on
    the same GPU; the cleaner with the oracle's one-thread Clean on the first
    4 views' depthmaps; accuracy with the oracle's on crops of 16 views, 2 seed sets."""
    from opensfm_amd import DepthmapBatch, DepthmapClusterEstimator, DepthmapEstimator, scenes
    from opensfm_amd._capi import PM_MODE_PATCHMATCH, TILE_NAMES
    log(f"batched workload: rendering {n_views} street views on the GPU")
    t0 = time.perf_counter()
    sc = scenes.street_scene(n_views, seed=seed, device="cuda")
    log(f"rendered in {time.perf_counter() - t0:.1f} s")
    npx = sum(sc.views[r].image.size for r, _ in sc.jobs)
    b = DepthmapBatch()
    for v in sc.views:
        b.add_view(v.K, v.R, v.t, v.image, v.mask)
    for ref, nb in sc.jobs:
        b.add_job(ref, nb, sc.min_depth, sc.max_depth)
    b.set_pyramid(BATCH_SCHEDULE[0])
    b.set_patchmatch_iterations(BATCH_SCHEDULE[1])
    b.set_pyramid_consistency(BATCH_SCHEDULE[2])
    b.set_min_patch_sd(1.0)
    b.set_seed(seed)

    def accuracy(batch):
        return float(np.mean([scenes.depth_accuracy(batch.result(k)[0], sc.views[r].depth,
                                                    sc.min_depth, sc.max_depth)
                              for k, (r, _) in enumerate(sc.jobs)]))

    t0 = time.perf_counter()
    patch, tile = b.autotune(patch_sizes=(5, 7), accuracy=accuracy, tolerance=0.002)
    tune_s = time.perf_counter() - t0
    log(f"autotune: patch {patch}, tile {TILE_NAMES[tile]} ({tune_s:.1f} s): {b.tune_report}")
    b.run()  # eager run that captures the graph
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run()
    batch_s = (time.perf_counter() - t0) / steps
    st = b.stats()
    accs, scores = [], []
    for k, (ref, _) in enumerate(sc.jobs):
        d, _, s, _ = b.result(k)
        accs.append(scenes.depth_accuracy(d, sc.views[ref].depth, sc.min_depth, sc.max_depth))
        scores.append(float(s[d > 0].mean()) if (d > 0).any() else 0.0)
    # plain PatchMatch at the reference's 3 iterations, same patch and tile
    b.set_pyramid(0)
    b.set_patchmatch_iterations(3)
    b.run()
    t0 = time.perf_counter()
    b.run()
    plain_s = time.perf_counter() - t0
    plain3 = {"iterations": 3, "mpix_s": round(npx / plain_s / 1e6, 3),
              "accuracy_within_1pct": round(accuracy(b), 4)}
    log(f"street batch, plain 3 iterations: {plain3}")
    b.set_pyramid(BATCH_SCHEDULE[0])
    b.set_patchmatch_iterations(BATCH_SCHEDULE[1])
    b.run()
    b.clean(0.01, 2)
    clean_ms = b.stats()["clean_ms"]
    kept, kept_good = 0, 0
    for k, (ref, _) in enumerate(sc.jobs):
        c = b.clean_result(k)
        gt = sc.views[ref].depth
        m = c > 0
        kept += int(m.sum())
        kept_good += int((m & (np.abs(c - gt) < 0.01 * gt)).sum())
    clean_cpu = None
    crop_ref = None
    if cpu_baseline:
        # DepthmapCleaner::Clean (depthmap.cc:524-555) by the oracle, one
        # thread, on the same packed depthmaps; the GPU's must be bit-equal
        import oracle
        depth = {ref: b.result(k)[0] for k, (ref, _) in enumerate(sc.jobs)}
        cpu_s, cpu_px, same, n_cmp = 0.0, 0, True, 0
        for k, (ref, nb) in enumerate(sc.jobs[:4]):
            if not all(v in depth for v in nb):
                continue
            views = [(sc.views[v].K, sc.views[v].R, sc.views[v].t, depth[v]) for v in [ref] + list(nb)]
            t0 = time.perf_counter()
            want = oracle.clean(views, 0.01, 2)
            cpu_s += time.perf_counter() - t0
            cpu_px += want.size
            n_cmp += 1
            same = same and bool(np.array_equal(want, b.clean_result(k)))
        if cpu_px:
            clean_cpu = {"views": n_cmp, "mpix_s": round(cpu_px / cpu_s / 1e6, 3), "cores": 1,
                         "kind": "port", "bit_exact": same}
        crop_ref = street_crops_oracle(sc, patch, 3, BATCH_SCHEDULE)
        log(f"street crops vs oracle: {crop_ref}")
    # one estimator per view, device-resident runs, same settings
    per_view_s = 0.0
    for k, (ref, nb) in enumerate(sc.jobs):
        e = DepthmapEstimator()
        for v in [ref] + list(nb):
            V = sc.views[v]
            e.add_view(V.K, V.R, V.t, V.image, V.mask)
        e.set_depth_range(sc.min_depth, sc.max_depth, 50)
        e.set_pyramid(BATCH_SCHEDULE[0])
        e.set_patchmatch_iterations(BATCH_SCHEDULE[1])
        e.set_pyramid_consistency(BATCH_SCHEDULE[2])
        e.set_patch_size(patch)
        e.set_min_patch_sd(1.0)
        e.set_seed(seed + k)
        t0 = time.perf_counter()
        e.run(PM_MODE_PATCHMATCH)
        per_view_s += time.perf_counter() - t0
    # the same jobs through the cluster estimator's surface (dense.py's call
    # sequence with its default depthmap_max_iterations = 4 and
    # depthmap_hierarchy_levels = 4, config.py:385, 411): host images in,
    # [depth, normal, cost, confidence] out
    ce = DepthmapClusterEstimator()
    ce.set_max_iterations(4)
    ce.set_hierarchy_levels(4)
    ce.set_patch_size(patch)
    for ref, nb in sc.jobs:
        R = sc.views[ref]
        ce.begin_ref_view(R.K, R.R, R.t, R.image, sc.min_depth, sc.max_depth)
        for v in nb:
            ce.add_source_view(sc.views[v].K, sc.views[v].R, sc.views[v].t, sc.views[v].image)
    ce.run()
    t0 = time.perf_counter()
    cres = ce.run()
    cluster_s = time.perf_counter() - t0
    cluster_acc = float(np.mean([scenes.depth_accuracy(r[0], sc.views[ref].depth, sc.min_depth,
                                                       sc.max_depth)
                                 for r, (ref, _) in zip(cres, sc.jobs)]))
    return {
        "views": n_views, "mpix": round(npx / 1e6, 2),
        "pyramid_levels": len(BATCH_SCHEDULE[0]) + 1,
        "coarse_level_iterations": list(BATCH_SCHEDULE[0]), "iterations": BATCH_SCHEDULE[1],
        "coarse_reselection": BATCH_SCHEDULE[2],
        "plain_reference_iterations": plain3,
        "patch": patch, "tile": TILE_NAMES[tile], "autotune_s": round(tune_s, 2),
        "autotune": b.tune_report,
        "mpix_s": round(npx / batch_s / 1e6, 3),
        "ms_per_run": round(batch_s * 1e3, 2), "device_ms": round(st["total_ms"], 2),
        # the batched red/black kernels' algorithmic rate over the whole run's
        # device time (every launch of the graph counted as red/black time,
        # so a lower bound on the kernels' own rate)
        "achieved_tflops_lower_bound": round(
            st["dominant_scorings"] * (FLOP_PER_SAMPLE * patch ** 2 + FLOP_PER_SCORING)
            / (st["total_ms"] / 1e3) / 1e12, 3),
        "frac_lower_bound": round(
            st["dominant_scorings"] * (FLOP_PER_SAMPLE * patch ** 2 + FLOP_PER_SCORING)
            / (st["total_ms"] / 1e3) / 1e12 / FP32_PEAK_TFLOPS, 4),
        "per_view_loop_mpix_s": round(npx / per_view_s / 1e6, 3),
        "accuracy_within_1pct": round(float(np.mean(accs)), 4),
        "mean_score": round(float(np.mean(scores)), 4),
        "oracle_crops": crop_ref,
        "clean_ms": round(clean_ms, 3),
        "clean_mpix_s": round(npx / (clean_ms / 1e3) / 1e6, 1) if clean_ms > 0 else None,
        "clean_kept_within_1pct": round(kept_good / max(1, kept), 4),
        "clean_cpu_baseline": clean_cpu,
        "cluster_estimator": {"max_iterations": 4, "hierarchy_levels": 4,
                              "wall_ms": round(cluster_s * 1e3, 1),
                              "mpix_s": round(npx / cluster_s / 1e6, 3),
                              "accuracy_within_1pct": round(cluster_acc, 4)},
    }


def precision_bench(steps=2, seed=3000):
    """Mixed-precision configuration: 16 reference views of 2048x1536 with 10
    neighbours each (the 1600x1200 scene's family on an arc of 26 cameras,
    rendered on the GPU), all in one batch, 5 iterations, with FP32, FP16 and
    BF16 samples (FP32 NCC sums, FP32 state, FP64 host homographies); then
    the same with the headline's pyramid (2, 2, 1 + 2 iterations)."""
    from opensfm_amd import DepthmapBatch, scenes
    log("mixed precision: rendering 26 views of 2048x1536 on the GPU")
    t0 = time.perf_counter()
    sc = scenes.scene_config3(seed=seed, device="cuda")
    log(f"rendered in {time.perf_counter() - t0:.1f} s")
    npx = sum(sc.views[r].image.size for r, _ in sc.jobs)
    b = DepthmapBatch()
    for v in sc.views:
        b.add_view(v.K, v.R, v.t, v.image, v.mask)
    for ref, nb in sc.jobs:
        b.add_job(ref, nb, sc.min_depth, sc.max_depth)
    b.set_patchmatch_iterations(5)
    b.set_patch_size(7)
    b.set_min_patch_sd(1.0)
    b.set_seed(seed)
    out = {"workload": "16 reference views 2048x1536, 10 neighbours each, patch 7, "
                       "5 iterations, one batch", "mpix": round(npx / 1e6, 2)}
    def timed_run():
        b.run()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.run()
        dt = (time.perf_counter() - t0) / steps
        accs, scores = [], []
        for k, (r, _) in enumerate(sc.jobs):
            d, _, s, _ = b.result(k)
            accs.append(scenes.depth_accuracy(d, sc.views[r].depth, sc.min_depth, sc.max_depth))
            scores.append(float(s[d > 0].mean()))
        return {"mpix_s": round(npx / dt / 1e6, 3), "ms_per_run": round(dt * 1e3, 1),
                "accuracy_within_1pct": round(float(np.mean(accs)), 4),
                "mean_score": round(float(np.mean(scores)), 4)}

    for prec in ("fp32", "fp16", "bf16"):
        b.set_precision(prec)
        out[prec] = timed_run()
        log(f"precision {prec}: {out[prec]}")
    b.set_pyramid(WORKLOAD["pyramid"])
    b.set_patchmatch_iterations(WORKLOAD["full_iterations"])
    b.set_pyramid_consistency(WORKLOAD["coarse_reselection"])
    out["pyramid"] = {"coarse_level_iterations": list(WORKLOAD["pyramid"]),
                      "iterations": WORKLOAD["full_iterations"]}
    for prec in ("fp32", "fp16", "bf16"):
        b.set_precision(prec)
        out["pyramid"][prec] = timed_run()
        log(f"precision {prec} with the pyramid: {out['pyramid'][prec]}")
    return out


def fusion_views(scene, seed=77):
    """The 1600x1200 scene's 9 views as fuser inputs: ground-truth depth with
    0.2 % noise and 2 % holes, normals from the ground truth's depth
    gradient, every view a neighbour (the pipeline's cluster lists)."""
    from opensfm_amd import scenes
    rng = np.random.default_rng(seed)
    n = len(scene.views)
    views = []
    for v in scene.views:
        d = scenes.noisy_depth(v, rng)
        views.append([v.K, v.R, v.t, d, scenes.normals_from_depth(v.K, v.depth),
                      np.stack([v.image] * 3, -1), None, list(range(n)), True])
    return views


def gpu_cleaner_bench(views, cpu_baseline=True):
    """GPUDepthmapCleaner (space carving + backward check, cleaner.cc:56-179)
    of view 0 against the other 8 views of the 1600x1200 scene (noisy
    ground-truth depths), beside the oracle's float32 restatement of the
    same clean on the host (one thread, numpy); the maps must be equal."""
    from opensfm_amd import GPUDepthmapCleaner
    c = GPUDepthmapCleaner()
    for v in views:
        c.add_view(v[0], v[1], v[2], v[3])
    nb = list(range(1, len(views)))
    c.clean(0, nb)  # warm-up
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        got = c.clean(0, nb)
    gpu_s = (time.perf_counter() - t0) / reps
    mpix = views[0][3].size / 1e6
    out = {"workload": f"view 0 of {len(views)} views {views[0][3].shape[1]}x{views[0][3].shape[0]} "
                       f"against {len(nb)} neighbours, defaults",
           "gpu_wall_ms": round(gpu_s * 1e3, 3), "mpix_s": round(mpix / gpu_s, 1),
           "kept": int((got > 0).sum())}
    if cpu_baseline:
        import oracle
        t0 = time.perf_counter()
        want = oracle.gpu_clean([tuple(v[:4]) for v in views], 0, nb)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"mpix_s": round(mpix / cpu_s, 3), "cores": 1, "kind": "port",
                               "bit_exact": bool(np.array_equal(got, want))}
        out["gpu_over_cpu"] = round(cpu_s / gpu_s, 1)
    return out


def fusion_bench(views):
    """The steps after estimation: DepthmapFuser over all views, 4 consistent
    views, margin 10, with and without outlier removal (k 12); then
    DepthmapPruner of view 0 against the others."""
    from opensfm_amd import DepthmapFuser, DepthmapPruner
    mpix = sum(v[3].size for v in views) / 1e6
    out = {"workload": f"{len(views)} views {views[0][3].shape[1]}x{views[0][3].shape[0]}, "
                       "all views neighbours, 4 consistent, margin 10"}
    for knn in (-1, 0, 12):           # -1: an untimed warm-up run
        f = DepthmapFuser()
        f.set_min_num_consistent(4)
        f.set_border_margin(10)
        f.set_sor_params(max(knn, 0), 2.5)
        for v in views:
            f.add_view(*v)
        t0 = time.perf_counter()
        pts = f.fuse()[0]
        wall = time.perf_counter() - t0
        st = f.stats()
        if knn < 0:
            continue
        out[f"gpu_sor{knn}"] = {"points": len(pts), "fused_before_sor": st["fused_before_sor"],
                                "rounds": st["rounds"], "fuse_ms": round(st["fuse_ms"], 2),
                                "sor_ms": round(st["sor_ms"], 2), "wall_ms": round(wall * 1e3, 2),
                                "mpix_s": round(mpix / wall, 2)}
    p = DepthmapPruner()
    for v in views:
        p.add_view(*v[:6], np.zeros(v[3].shape, np.uint8))
    p.prune()
    t0 = time.perf_counter()
    kept = len(p.prune()[0])
    out["prune"] = {"kept": kept, "gpu_ms": round(p.last_ms(), 3),
                    "gpu_wall_ms": round((time.perf_counter() - t0) * 1e3, 2)}
    return out


def fusion_cpu_baseline(views, primary=2):
    """CPU baseline of the fusion section: the oracle's one-thread fuser over
    the first `primary` views (all views as sources) and the oracle pruner
    of view 0, on this host."""
    import oracle
    sub = [v[:8] + [k < primary] for k, v in enumerate(views)]
    t0 = time.perf_counter()
    _, nf = oracle.fuse(sub, min_num_consistent=4, border_margin=10, sor_knn=0)
    fuse_s = time.perf_counter() - t0
    mpix = sum(v[3].size for v in views[:primary]) / 1e6
    pv = [tuple(v[:6]) + (np.zeros(v[3].shape, np.uint8),) for v in views]
    t0 = time.perf_counter()
    kept = len(oracle.prune(pv)[0])
    prune_s = time.perf_counter() - t0
    return {"fuse": {"primary_views": primary, "points": nf, "s": round(fuse_s, 2),
                     "mpix_s": round(mpix / fuse_s, 3), "cores": 1, "kind": "port"},
            "prune": {"kept": kept, "ms": round(prune_s * 1e3, 1), "cores": 1, "kind": "port"}}


def neighbors_bench(nx=25, ny=20, n_points=300000):
    """Neighbour selection before the depthmaps on a synthetic sparse
    reconstruction (nx x ny cameras, n_points landmarks on a wavy ground,
    each seen by the cameras within 25 m): covisibility graph (k 15, radius
    factor 0.5) and 10 neighbours per shot, on the GPU: device time between
    the first upload and the last kernel, and wall time including the
    Python objects of the result."""
    from opensfm_amd import neighbors, scenes
    origins, coords, obs = scenes.sfm_landmarks(nx, ny, n_points)
    ids = list(range(len(origins)))
    off = np.zeros(n_points + 1, np.int64)
    off[1:] = np.cumsum([len(o) for o in obs])
    obs = (off, np.concatenate([np.asarray(o, np.int32) for o in obs]))   # CSR form
    neighbors.build_covisibility_graph(coords[:2000], (off[:2001], obs[1][:off[2000]]), ids)
    t0 = time.perf_counter()
    g = neighbors.build_covisibility_graph(coords, obs, ids)
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    r = neighbors.select_neighbors(origins, g.super_points, ids, 10)
    select_s = time.perf_counter() - t0
    nobs = int(off[-1])
    return {"workload": f"{len(ids)} shots, {n_points} landmarks, {nobs} observations",
            "super_points": len(g.super_points), "edges": len(g.edges),
            "build_device_ms": round(g.device_ms, 2), "select_device_ms": round(r.device_ms, 2),
            "build_wall_ms": round(build_s * 1e3, 1), "select_wall_ms": round(select_s * 1e3, 1),
            "mean_best": round(float(np.mean([len(v) for v in r.best_neighbors.values()])), 2),
            "mpts_s": round(n_points / (g.device_ms / 1e3) / 1e6, 3)}


def neighbors_cpu_baseline(n_points=20000, nx=25, ny=20):
    """CPU baseline of the neighbour section: the oracle's graph (brute-force
    k nearest) on a bounded sample of the same reconstruction, one thread."""
    import oracle
    from opensfm_amd import scenes
    origins, coords, obs = scenes.sfm_landmarks(nx, ny, n_points)
    t0 = time.perf_counter()
    o = oracle.covis(coords, obs, len(origins))
    s = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.select_neighbors(origins, o["sp_coord"], o["sp_vis"],
                            [len(m) for m in o["sp_members"]], 10)
    return {"sample_points": n_points, "build_s": round(s, 2),
            "select_s": round(time.perf_counter() - t0, 2),
            "mpts_s": round(n_points / s / 1e6, 4), "cores": 1, "kind": "port"}


def learned_bench(rows=8, cols=8, width=4000, height=3000, seed=5000):
    """Learned initialiser configuration: 64 nadir views of 4000x3000 over
    fractal terrain (an 8 x 8 camera grid, rendered on the GPU), 10
    neighbours each, all in one DepthmapBatch.  Plain PatchMatch from random
    planes at 5, 2 and 1 iterations, and through a 4-level pyramid (2, 2, 2
    + 2 and 2, 2, 1 + 1 iterations), against the learned initialiser
    (network input from the batch's 1/4-resolution views and 32-plane cost
    volume, the U-Net, then seeded PatchMatch) at 1 and 2 iterations, and
    with its map seeding the 1/4 level of a 3-level pyramid (1, 1 + 1
    iterations), each timed once from device-resident views to
    device-resident maps, the network inside the timed region."""
    import torch
    from opensfm_amd import DepthmapBatch, initializer as I, nn, scenes
    log(f"learned initialiser: rendering {rows * cols} nadir views of {width}x{height} on the GPU")
    t0 = time.perf_counter()
    sc = scenes.nadir_scene(rows, cols, width, height, seed=seed, neighbours=10, device="cuda")
    log(f"rendered in {time.perf_counter() - t0:.1f} s")
    jobs = sc.jobs
    npx = sum(sc.views[r].image.size for r, _ in jobs)
    b = DepthmapBatch()
    for v in sc.views:
        b.add_view(v.K, v.R, v.t, v.image, v.mask)
    for r, nb in jobs:
        b.add_job(r, nb, sc.min_depth, sc.max_depth)
    b.set_patch_size(7)
    b.set_min_patch_sd(1.0)
    b.set_seed(seed)
    b.set_graph(False)
    init = I.LearnedInitializer(weights=I.WEIGHTS)

    def accuracy():
        return float(np.mean([scenes.depth_accuracy(b.result(k)[0], sc.views[r].depth,
                                                    sc.min_depth, sc.max_depth)
                              for k, (r, _) in enumerate(jobs)]))

    out = {"workload": f"{len(jobs)} nadir views {width}x{height}, 10 neighbours, patch 7, one "
                       f"batch", "mpix": round(npx / 1e6, 1), "params": init.net.n_params}
    # warm the network and the coarse views outside the timed region
    init.predict_batch(b, 0, width, height)
    # coarser levels' iterations, coarsest first (None: one level); the
    # learned map seeds the coarsest level when there is a pyramid
    schedules = {"plain_5": (False, 5, None), "plain_2": (False, 2, None),
                 "plain_1": (False, 1, None), "plain_pyramid_2": (False, 2, (2, 2, 2)),
                 "plain_pyramid_1": (False, 1, (2, 2, 1)), "learned_1": (True, 1, None),
                 "learned_2": (True, 2, None), "learned_pyramid_1": (True, 1, (1, 1))}
    out["schedules"] = {m: {"coarse_level_iterations": list(p) if p else None,
                            "full_iterations": it, "learned_map": lrn}
                        for m, (lrn, it, p) in schedules.items()}
    for mode, (learned, iters, pyr) in schedules.items():
        b.set_pyramid(pyr if pyr else 0)
        b.set_patchmatch_iterations(iters)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if learned:
            for k, (r, _) in enumerate(jobs):
                m, ow, oh = init.predict_batch(b, k, width, height)
                b.set_init_map(k, m, ow, oh, nn.OUT_CHANNELS, I.SCALE)
            torch.cuda.synchronize()
        t_net = time.perf_counter() - t0
        b.run()
        dt = time.perf_counter() - t0
        out[mode] = {"acc": round(accuracy(), 4), "s": round(dt, 2),
                     "mpix_s": round(npx / dt / 1e6, 2)}
        if learned:
            out[mode]["net_s"] = round(t_net, 2)
            for k in range(len(jobs)):
                b.set_init_map(k, None)
        log(f"learned initialiser {mode}: {out[mode]}")
    base = out["plain_5"]["mpix_s"]
    for mode in schedules:
        if mode != "plain_5":
            out[f"{mode}_over_plain_5"] = round(out[mode]["mpix_s"] / base, 2)
    # the fastest plain schedule reaching plain-5's accuracy, and the
    # learned runs at that accuracy
    target = out["plain_5"]["acc"]
    eq_plain = max((m for m in schedules if m.startswith("plain") and out[m]["acc"] >= target),
                   key=lambda m: out[m]["mpix_s"])
    eq_learned = sorted((m for m in schedules
                         if m.startswith("learned") and out[m]["acc"] >= target),
                        key=lambda m: -out[m]["mpix_s"])
    out["equal_accuracy"] = {
        "plain_fastest_at_plain_5_accuracy": eq_plain,
        "plain_fastest_over_plain_5": round(out[eq_plain]["mpix_s"] / base, 2),
        "learned_at_equal_accuracy": eq_learned[0] if eq_learned else None,
        "ratio_vs_plain_5": (round(out[eq_learned[0]]["mpix_s"] / base, 2) if eq_learned else None),
        "ratio_vs_fastest_plain": (round(out[eq_learned[0]]["mpix_s"] / out[eq_plain]["mpix_s"], 2)
                                   if eq_learned else None)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--batch-views", type=int, default=128, metavar="N",
                    help="also run the batched street workload with N reference "
                         "views (0: skip; 128 is the full batched configuration)")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the mixed-precision section (16 views of 2048x1536)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-initialiser section (64 nadir views of 4000x3000)")
    ap.add_argument("--no-fusion", action="store_true",
                    help="skip the GPU cleaner / fusion / pruning sections")
    ap.add_argument("--no-neighbors", action="store_true",
                    help="skip the neighbour-selection section")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run one plain device copy of MIB MiB (torch) so that "
                         "profiler byte counters can be scaled against a known size")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the depthmap path shards by reference view, one process per GPU)",
              file=sys.stderr)
        sys.exit(2)

    w = WORKLOAD
    log("rendering the 1600x1200 scene (9 views)")
    scene = make_scene()

    cpu = None
    crops = []
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or max(1, min(16, len(os.sched_getaffinity(0))))
        log(f"CPU baseline: oracle PatchMatch on {threads} threads")
        crops, cpu = cpu_baseline(scene, threads)
        log(f"CPU baseline {cpu['value']:.4f} Mpix/s ({cpu['wall_s']:.1f} s)")

    from opensfm_amd import DepthmapEstimator, scenes
    from opensfm_amd._capi import PM_MODE_PATCHMATCH, PM_MODE_PATCHMATCH_SAMPLE, PM_MODE_PLANE_SWEEP

    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        # one elementwise kernel that reads and writes exactly MIB MiB each
        # (a same-dtype copy_ would go to the DMA engine, not a kernel)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calibration} MiB read + written")
        del src, dst

    est = DepthmapEstimator()
    for v in scene.views:
        est.add_view(v.K, v.R, v.t, v.image, v.mask)
    est.set_depth_range(w["min_depth"], w["max_depth"], 64)
    est.set_patch_size(w["patch"])
    est.set_min_patch_sd(w["min_patch_sd"])
    est.set_seed(w["seed"])

    def schedule(half):
        """half: the pyramid (coarser levels + full iterations) that reaches
        the reference's 5-iteration accuracy; else plain PatchMatch at the
        reference's 5 iterations."""
        est.set_pyramid(w["pyramid"] if half else 0)
        est.set_pyramid_consistency(w["coarse_reselection"])
        est.set_patchmatch_iterations(w["full_iterations"] if half else w["iterations"])

    def timed(steps, warmup):
        for _ in range(warmup):
            est.run(PM_MODE_PATCHMATCH)
        dom = [0.0, 0, 0]
        t0 = time.perf_counter()
        for _ in range(steps):
            est.run(PM_MODE_PATCHMATCH)  # synchronous: returns after the stream drains
            st = est.stats()
            dom[0] += st["dominant_ms"]
            dom[1] += st["launches"]
            dom[2] += st["dominant_scorings"]
        return time.perf_counter() - t0, dom

    schedule(True)
    log(f"warmup {args.warmup}, {args.steps} timed steps")
    elapsed, (dom_ms, dom_launches, dom_scorings) = timed(args.steps, args.warmup)
    log(f"timed {args.steps} steps in {elapsed:.3f} s")
    depth, plane, score, nghbr = est.download()

    # PCIe-inclusive rate: compute + download of the four maps
    t1 = time.perf_counter()
    est.compute_patch_match()
    pcie_s = time.perf_counter() - t1

    # equal accuracy: the same compute over seeds 1-5 against the oracle's
    # full-view figures for those seeds (recorded by
    # tests/golden/make_config2_oracle.py, about 1300 s per seed on one core)
    with open(GOLDEN_CFG2) as f:
        fix = json.load(f)

    def against_oracle():
        g_acc, g_score = [], []
        for r in fix["per_seed"]:
            est.set_seed(r["seed"])
            est.run(PM_MODE_PATCHMATCH)
            d_s, _, s_s, _ = est.download()
            g_acc.append(scenes.depth_accuracy(d_s, scene.ref.depth, w["min_depth"],
                                               w["max_depth"]))
            g_score.append(float(s_s[d_s > 0].mean()))
        est.set_seed(w["seed"])
        eq = {"seeds": [r["seed"] for r in fix["per_seed"]],
              "oracle_accuracy": round(fix["mean_accuracy_within_1pct"], 5),
              "oracle_mean_score": round(fix["mean_score"], 5),
              "gpu_accuracy": round(float(np.mean(g_acc)), 5),
              "gpu_mean_score": round(float(np.mean(g_score)), 5)}
        eq["met"] = bool(eq["gpu_accuracy"] >= eq["oracle_accuracy"] and
                         eq["gpu_mean_score"] >= eq["oracle_mean_score"])
        return eq

    equal = against_oracle()
    log(f"equal accuracy (full view, 5 seeds): {equal}")

    # plain PatchMatch at the reference's 5 iterations, beside it
    schedule(False)
    p_elapsed, _ = timed(max(1, min(args.steps, 3)), 1)
    p_depth = est.download()[0]
    plain = {"iterations": w["iterations"],
             "mpix_s": round(w["width"] * w["height"] * max(1, min(args.steps, 3)) / p_elapsed
                             / 1e6, 3),
             "accuracy_within_1pct": round(scenes.depth_accuracy(p_depth, scene.ref.depth,
                                                                 w["min_depth"], w["max_depth"]), 4),
             "equal_accuracy": against_oracle()}
    log(f"plain PatchMatch, 5 iterations: {plain}")

    # ComputePlaneSweep over 64 fronto-parallel planes (exact FP32 path)
    est.run(PM_MODE_PLANE_SWEEP)
    t2 = time.perf_counter()
    est.run(PM_MODE_PLANE_SWEEP)
    sweep_s = time.perf_counter() - t2
    sst = est.stats()
    sweep_flop = sst["dominant_scorings"] * (FLOP_PER_SAMPLE * w["patch"] ** 2 + FLOP_PER_SCORING)
    sweep = {"planes": 64, "mpix_s": round(w["width"] * w["height"] / sweep_s / 1e6, 3),
             "kernel_ms": round(sst["dominant_ms"], 3),
             "achieved_tflops": round(sweep_flop / (sst["dominant_ms"] / 1e3) / 1e12, 3)}
    log(f"plane sweep 64 planes: {sweep}")

    # The reference's iteration is a forward and a backward sweep (each pixel
    # updated twice, depthmap.cc:207-210); a red/black iteration updates each
    # pixel once.  The same number of pixel updates: 2 x iterations.
    est.set_patchmatch_iterations(2 * w["iterations"])
    est.run(PM_MODE_PATCHMATCH)
    t3 = time.perf_counter()
    est.run(PM_MODE_PATCHMATCH)
    upd_s = time.perf_counter() - t3
    upd_depth = est.download()[0]
    matched_updates = {
        "iterations": 2 * w["iterations"],
        "mpix_s": round(w["width"] * w["height"] / upd_s / 1e6, 3),
        "accuracy_within_1pct": round(scenes.depth_accuracy(upd_depth, scene.ref.depth,
                                                            w["min_depth"], w["max_depth"]), 4)}
    log(f"matched pixel updates: {matched_updates}")

    # ComputePatchMatchSample (depthmap.cc:215-227): each candidate scored
    # against one view, at the reference's iterations on one level, and with
    # the headline's pyramid (every level sampling)
    def sample_run():
        est.run(PM_MODE_PATCHMATCH_SAMPLE)
        t5 = time.perf_counter()
        est.run(PM_MODE_PATCHMATCH_SAMPLE)
        smp_s = time.perf_counter() - t5
        smp_depth = est.download()[0]
        return {"mpix_s": round(w["width"] * w["height"] / smp_s / 1e6, 3),
                "accuracy_within_1pct": round(scenes.depth_accuracy(
                    smp_depth, scene.ref.depth, w["min_depth"], w["max_depth"]), 4)}

    schedule(False)
    sample_mode = {"iterations": w["iterations"], **sample_run()}
    schedule(True)
    sample_mode["pyramid"] = {"pyramid_levels": len(w["pyramid"]) + 1,
                              "coarse_level_iterations": list(w["pyramid"]),
                              "iterations": w["full_iterations"], **sample_run()}
    log(f"sample mode: {sample_mode}")

    npx = w["width"] * w["height"]
    ms_per_step = elapsed / args.steps * 1e3
    value = npx * args.steps / elapsed / 1e6
    acc = scenes.depth_accuracy(depth, scene.ref.depth, w["min_depth"], w["max_depth"])
    mean_score = float(score[depth > 0].mean()) if (depth > 0).any() else 0.0

    p2 = w["patch"] ** 2
    flop_per_launch = dom_scorings / max(1, dom_launches) * (FLOP_PER_SAMPLE * p2 + FLOP_PER_SCORING)
    avg_launch_s = dom_ms / max(1, dom_launches) / 1e3
    achieved = flop_per_launch / avg_launch_s / 1e12 if avg_launch_s > 0 else 0.0
    out = {
        "metric": "PatchMatch Mpix/s at equal accuracy (\u0394Z<1%); learned-init vs plain, same GPU",
        "value": round(value, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ray-cast scene, value-noise texture)",
        "config": {"workload": "single-view plain PatchMatch 1600x1200 + 8 neighbours, "
                               "patch 7, depth [1, 40], FP32, at the accuracy and mean score of the "
                               "reference's 5 iterations: a 4-level pyramid (1/8, 1/4, 1/2, full "
                               "resolution) with 2, 2, 1 and 2 iterations, each level seeding the "
                               "next, the re-selection at the full level only; "
                               "value = device-resident views to device-resident maps",
                   "width": w["width"], "height": w["height"], "neighbours": w["neighbours"],
                   "patch": w["patch"], "pyramid_levels": len(w["pyramid"]) + 1,
                   "coarse_level_iterations": list(w["pyramid"]),
                   "coarse_reselection": w["coarse_reselection"],
                   "iterations": w["full_iterations"],
                   "reference_iterations": w["iterations"]},
        "accuracy_within_1pct": round(acc, 4),
        "equal_accuracy": equal,
        "plain_reference_iterations": plain,
        "matched_pixel_updates": matched_updates,
        "sample_mode": sample_mode,
        "mean_score": round(mean_score, 4),
        "pcie_inclusive_mpix_s": round(npx / pcie_s / 1e6, 3),
        "plane_sweep": sweep,
        "roofline": {
            "kernel": "k_pm_half<7> (red/black PatchMatch half-iteration, every launch of a "
                      "compute at all pyramid levels: k_pm_half_group<7> with 8 lanes per pixel "
                      "on the 200x150 and 400x300 levels, k_pm_half<7> on the others; "
                      "random-plane launches read the u8 quads, the others the FP16 "
                      "coefficients)",
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "traffic": pmc_traffic(),
            "traffic_source": "rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 per launch, "
                              "profiles/pm_half_p7_1600x1200_pmc.md",
            "avg_launch_ms": round(avg_launch_s * 1e3, 4),
            "scorings_per_launch": int(dom_scorings / max(1, dom_launches)),
        },
        "cpu_baseline": None if cpu is None else {k: (round(v, 6) if isinstance(v, float) else v)
                                                   for k, v in cpu.items() if k != "wall_s"},
    }
    if cpu is not None:
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
        # equal-accuracy check on identical inputs: the GPU engine on the
        # CPU baseline's crops (same views, K shift, depth range, iterations)
        crop_est = []
        for k, (x0, y0, cw, ch, odepth) in enumerate(crops):
            ce = DepthmapEstimator()
            for K, R, t, img, mask in crop_views(scene, x0, y0, cw, ch):
                ce.add_view(K, R, t, img, mask)
            ce.set_depth_range(w["min_depth"], w["max_depth"], 50)
            ce.set_patch_size(w["patch"])
            ce.set_min_patch_sd(w["min_patch_sd"])
            ce.set_seed(100 + k)
            crop_est.append(ce)

        def gpu_crops():
            acc, inner = [], []   # inner: at least 20 px from the crop's edges
            for ce, (x0, y0, cw, ch, _) in zip(crop_est, crops):
                gt = scene.ref.depth[y0:y0 + ch, x0:x0 + cw]
                ce.set_pyramid(w["pyramid"])
                ce.set_pyramid_consistency(w["coarse_reselection"])
                ce.set_patchmatch_iterations(w["full_iterations"])
                gd = ce.compute_patch_match()[0]
                acc.append(scenes.depth_accuracy(gd, gt, w["min_depth"], w["max_depth"]))
                inner.append(scenes.depth_accuracy(gd[20:-20, 20:-20], gt[20:-20, 20:-20],
                                                   w["min_depth"], w["max_depth"]))
            return float(np.mean(acc)), float(np.mean(inner))

        cpu_acc, cpu_in = [], []
        for x0, y0, cw, ch, odepth in crops:
            gt = scene.ref.depth[y0:y0 + ch, x0:x0 + cw]
            cpu_acc.append(scenes.depth_accuracy(odepth, gt, w["min_depth"], w["max_depth"]))
            cpu_in.append(scenes.depth_accuracy(odepth[20:-20, 20:-20], gt[20:-20, 20:-20],
                                                w["min_depth"], w["max_depth"]))
        gpu_mean, gpu_inner = gpu_crops()
        # the CPU baseline's own crops (seed set 0) through the GPU engine:
        # crops have a larger share of image-edge pixels than the full view
        out["crop_check"] = {
            "crops": len(crops),
            "cpu_reference_accuracy": round(float(np.mean(cpu_acc)), 4),
            "gpu_accuracy": round(gpu_mean, 4),
            "cpu_reference_accuracy_interior": round(float(np.mean(cpu_in)), 4),
            "gpu_accuracy_interior": round(gpu_inner, 4),
        }
    if args.batch_views > 0:
        out["batched"] = batch_bench(args.batch_views, cpu_baseline=not args.no_cpu_baseline)
        log(f"batched: {out['batched']}")
    if not args.no_precision:
        out["precision"] = precision_bench()
    if not args.no_fusion:
        fv = fusion_views(scene)
        out["gpu_cleaner"] = gpu_cleaner_bench(fv, cpu_baseline=not args.no_cpu_baseline)
        log(f"gpu cleaner: {out['gpu_cleaner']}")
        out["fusion"] = fusion_bench(fv)
        if not args.no_cpu_baseline:
            out["fusion"]["cpu_baseline"] = fusion_cpu_baseline(fv)
            out["fusion"]["fuse_gpu_over_cpu"] = round(
                out["fusion"]["gpu_sor0"]["mpix_s"] / out["fusion"]["cpu_baseline"]["fuse"]["mpix_s"], 1)
        log(f"fusion: {out['fusion']}")
    if not args.no_neighbors:
        out["neighbors"] = neighbors_bench()
        if not args.no_cpu_baseline:
            out["neighbors"]["cpu_baseline"] = neighbors_cpu_baseline()
        log(f"neighbors: {out['neighbors']}")
    if not args.no_learned:
        out["learned_init"] = learned_bench()
        log(f"learned initialiser: {out['learned_init']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
