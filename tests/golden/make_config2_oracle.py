"""Writes tests/golden/config2_oracle_fullview.json: the oracle's full-view
ComputePatchMatch accuracy and mean score on the 1600x1200 + 8-neighbour
scene (20 boxes on a sinusoidal terrain, depth [1, 40], patch 7, min patch
SD 1, 5 iterations), one entry per RNG seed.

The oracle restates opensfm/src/lib/dense/src/depthmap.cc:202-213 (its
sequential forward/backward sweeps, depthmap.cc:283-374) and takes about
1000 s per seed on one core, so the GPU tests compare against these
recorded figures instead of running it.  Seeds run concurrently, one host
thread each (ctypes releases the GIL).

    python tests/golden/make_config2_oracle.py [--seeds 1 2 3 4 5] [--dump DIR]

--dump DIR also saves each seed's four maps as DIR/oracle_seed<k>.npz.
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
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from opensfm_amd import scenes  # noqa: E402

OUT = os.path.join(HERE, "config2_oracle_fullview.json")
WORKLOAD = dict(scene_seed=2000, width=1600, height=1200, neighbours=8, min_depth=1.0,
                max_depth=40.0, patch=7, min_patch_sd=1.0, iterations=5)


def run_seed(scene, seed, dump):
    w = WORKLOAD
    est = oracle.OracleEstimator()
    for v in scene.views:
        est.add_view(v.K, v.R, v.t, v.image, v.mask)
    est.set_depth_range(w["min_depth"], w["max_depth"], 50)
    est.set_patch_size(w["patch"])
    est.set_min_patch_sd(w["min_patch_sd"])
    est.set_patchmatch_iterations(w["iterations"])
    est.set_seed(seed)
    t0 = time.perf_counter()
    depth, plane, score, nghbr = est.compute_patch_match()
    dt = time.perf_counter() - t0
    gt = scene.views[0].depth
    live = depth > 0
    res = dict(seed=seed, seconds=round(dt, 1),
               accuracy_within_1pct=scenes.depth_accuracy(depth, gt),
               mean_score=float(score[live].mean()),
               live_pixels=int(live.sum()), scorings=int(est.last_scorings))
    if dump:
        np.savez_compressed(os.path.join(dump, f"oracle_seed{seed}.npz"), depth=depth,
                            plane=plane, score=score, nghbr=nghbr)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, nargs="+", default=[1, 2, 3, 4, 5])
    ap.add_argument("--dump", default=None)
    args = ap.parse_args()
    if args.dump:
        os.makedirs(args.dump, exist_ok=True)
    scene = scenes.scene_config2(seed=WORKLOAD["scene_seed"], width=WORKLOAD["width"],
                                 height=WORKLOAD["height"])
    results = [None] * len(args.seeds)

    def worker(k):
        results[k] = run_seed(scene, args.seeds[k], args.dump)
        print(json.dumps(results[k]), flush=True)

    ths = [threading.Thread(target=worker, args=(k,)) for k in range(len(args.seeds))]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    acc = [r["accuracy_within_1pct"] for r in results]
    sc = [r["mean_score"] for r in results]
    doc = dict(workload=WORKLOAD, per_seed=results,
               mean_accuracy_within_1pct=float(np.mean(acc)),
               mean_score=float(np.mean(sc)))
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
    print(json.dumps({k: doc[k] for k in ("mean_accuracy_within_1pct", "mean_score")}))


if __name__ == "__main__":
    main()
