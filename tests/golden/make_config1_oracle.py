"""Writes tests/golden/config1_oracle.json: the oracle's ComputePatchMatch and
ComputePatchMatchSample accuracy and mean score on the 320x240 + 4-neighbour
scene (ground plane and 6 boxes, depth [2, 20], patch 7, min patch SD 1,
3 iterations), seeds 1-4, so that the GPU tests compare against recorded
figures instead of running the oracle (opensfm/src/lib/dense/src/
depthmap.cc:202-227) on the GPU box's host.

    python tests/golden/make_config1_oracle.py
"""
from __future__ import annotations

import json
import os
import sys
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from opensfm_amd import scenes  # noqa: E402

OUT = os.path.join(HERE, "config1_oracle.json")
SEEDS = (1, 2, 3, 4)


def run(scene, seed, sample):
    est = oracle.OracleEstimator()
    for v in scene.views:
        est.add_view(v.K, v.R, v.t, v.image, v.mask)
    est.set_depth_range(scene.min_depth, scene.max_depth, 16)
    est.set_patch_size(7)
    est.set_min_patch_sd(1.0)
    est.set_patchmatch_iterations(3)
    est.set_seed(seed)
    d, _, s, _ = est.compute_patch_match_sample() if sample else est.compute_patch_match()
    return dict(seed=seed, accuracy_within_1pct=scenes.depth_accuracy(d, scene.ref.depth, 2.0, 20.0),
                mean_score=float(s[d > 0].mean()))


def main():
    scene = scenes.scene_config1()
    res = {}
    jobs = [(m, s) for m in ("patchmatch", "patchmatch_sample") for s in SEEDS]
    out = [None] * len(jobs)

    def worker(k):
        m, s = jobs[k]
        out[k] = run(scene, s, m == "patchmatch_sample")

    ths = [threading.Thread(target=worker, args=(k,)) for k in range(len(jobs))]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    for (m, _), r in zip(jobs, out):
        res.setdefault(m, {"per_seed": []})["per_seed"].append(r)
    for m, d in res.items():
        d["mean_accuracy_within_1pct"] = float(np.mean([r["accuracy_within_1pct"] for r in d["per_seed"]]))
        d["mean_score"] = float(np.mean([r["mean_score"] for r in d["per_seed"]]))
    doc = {"workload": "config 1: 320x240 + 4 neighbours, 6 boxes, depth [2, 20], patch 7, "
                       "min patch SD 1, 3 iterations, scene seed 1000", **res}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
    print(json.dumps({m: (d["mean_accuracy_within_1pct"], d["mean_score"]) for m, d in res.items()}))


if __name__ == "__main__":
    main()
