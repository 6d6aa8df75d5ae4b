"""Reduced-precision samples (FP16 / BF16) against an FP64 recomputation.

The PatchMatch scorer holds image samples as FP16 or BF16 2x2 quads and
blends each row with one FP16/BF16 dot product accumulated in FP32; the NCC
sums stay FP32.  Its error is stated against the oracle's FP64 twin of
ComputePlaneImageScore (depthmap.cc:441-478 with every quantity in double)
on 10^5 random (pixel, hypothesis) pairs: uniform interior pixels, depth
log-uniform in the depth range, normal (U(-1,1), U(-1,1), -1), a random
neighbour view.  Bounds (written here, measured on MI355X with margin):
pairs where both scores are defined (not -1) must agree to
  FP32 fast scorer: max 5e-4,  FP16: max 5e-3,  BF16: max 4e-2
(measured: 1.5e-4, 1.5e-3, 1.3e-2; means 6e-6, 6.5e-5, 5.3e-4)
and full PatchMatch at FP16/BF16 must keep accuracy and mean score within
0.005 of the FP32 run.  Both checks run at 640x480 and at the mixed-precision
configuration's size (16 reference views of 2048x1536, 10 neighbours each).
"""
import numpy as np
import pytest

import oracle
from opensfm_amd import DepthmapBatch, DepthmapEstimator, scenes

pytestmark = pytest.mark.gpu

N_PAIRS = 100_000
BOUNDS = {"fp32": 5e-4, "fp16": 5e-3, "bf16": 4e-2}


@pytest.fixture(scope="module")
def wide_scene():
    # the single-view scene family (boxes on sinusoidal terrain) with 10
    # neighbours at 4 degree spacing
    return scenes.make_scene(640, 480, 10, 8, seed=3000, min_depth=1.0, max_depth=40.0,
                             terrain=(1.5, 6.0), spacing_deg=4.0)


def _add(est, scene, patch=7, iters=3):
    for v in scene.views:
        est.add_view(v.K, v.R, v.t, v.image, v.mask)
    est.set_depth_range(scene.min_depth, scene.max_depth, 50)
    est.set_patch_size(patch)
    est.set_min_patch_sd(1.0)
    est.set_patchmatch_iterations(iters)
    return est


def _pairs(scene, n, seed=12):
    rng = np.random.default_rng(seed)
    H, W = scene.ref.image.shape
    i = rng.integers(3, H - 3, n)
    j = rng.integers(3, W - 3, n)
    depth = np.exp(rng.uniform(np.log(scene.min_depth), np.log(scene.max_depth), n))
    normals = np.stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), -np.ones(n)], 1)
    K = scene.ref.K
    planes = np.stack([oracle.plane_from_depth_and_normal(int(j[k]), int(i[k]), K,
                                                          np.float32(depth[k]),
                                                          normals[k].astype(np.float32))
                       for k in range(n)]).astype(np.float32)
    views = rng.integers(1, len(scene.views), n).astype(np.int32)
    return np.stack([i, j], 1).astype(np.int32), planes, views


def test_sample_precision_error_vs_fp64(wide_scene):
    print("scorer error vs FP64 on 1e5 pairs, 640x480 + 10 neighbours", flush=True)
    ij, planes, views = _pairs(wide_scene, N_PAIRS)
    o = _add(oracle.OracleEstimator(), wide_scene)
    ref = np.array([o.score_f64(int(a), int(b), p.astype(np.float64), int(v))
                    for (a, b), p, v in zip(ij, planes, views)])
    e = _add(DepthmapEstimator(), wide_scene)
    exact = e.score_hypotheses(ij, planes, views)
    report = {}
    for prec in ("fp32", "fp16", "bf16"):
        e.set_precision(prec)
        got = e.score_hypotheses(ij, planes, views, fast=True).astype(np.float64)
        # defined in both (a pair on the -1 rule's cut-off counts as a flip)
        ok = (ref > -1.0) & (got > -1.0)
        err = np.abs(got - ref)[ok]
        flips = int(((ref > -1.0) != (got > -1.0)).sum())
        report[prec] = (float(err.max()), float(np.percentile(err, 99.9)), float(err.mean()), flips)
        print(f"{prec}: max {err.max():.3e} p99.9 {np.percentile(err, 99.9):.3e} "
              f"mean {err.mean():.3e} over {ok.sum()} pairs, -1 flips {flips}", flush=True)
        assert ok.sum() > 0.3 * N_PAIRS
        assert err.max() <= BOUNDS[prec], (prec, err.max())
        assert flips <= 0.002 * N_PAIRS
    ok = (ref > -1.0) & (exact > -1.0)
    print(f"exact FP32 scorer: max {np.abs(exact - ref)[ok].max():.3e}", flush=True)
    assert np.abs(exact - ref)[ok].max() <= 5e-4


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_patchmatch_reduced_precision_accuracy(cfg1_scene, prec):
    print(f"PatchMatch config 1 at {prec}", flush=True)
    gt = cfg1_scene.ref.depth
    res = {}
    for p in ("fp32", prec):
        accs, scores = [], []
        for seed in (5, 6):
            e = _add(DepthmapEstimator(), cfg1_scene)
            e.set_precision(p)
            e.set_seed(seed)
            d, _, s, _ = e.compute_patch_match()
            accs.append(scenes.depth_accuracy(d, gt, 2.0, 20.0))
            scores.append(float(s[d > 0].mean()))
        res[p] = (np.mean(accs), np.mean(scores))
    print(f"fp32 acc {res['fp32'][0]:.4f} score {res['fp32'][1]:.4f} | "
          f"{prec} acc {res[prec][0]:.4f} score {res[prec][1]:.4f}", flush=True)
    assert res[prec][0] >= res["fp32"][0] - 0.005
    assert res[prec][1] >= res["fp32"][1] - 0.005


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_reduced_precision_with_pyramid(wide_scene, prec):
    """The pyramid at FP16/BF16 (every level builds its 16-bit quads; lane
    groups are FP32-only, so every level runs one lane per pixel): the
    batch equals the single-view estimator bit for bit, and the estimator's
    accuracy and mean score stay within 0.005 of its FP32 pyramid run."""
    print(f"pyramid at {prec}", flush=True)
    gt = wide_scene.ref.depth
    res = {}
    for p in ("fp32", prec):
        e = _add(DepthmapEstimator(), wide_scene, iters=2)
        e.set_precision(p)
        e.set_seed(4)
        e.set_pyramid((2, 1))
        d, _, s, _ = e.compute_patch_match()
        res[p] = (scenes.depth_accuracy(d, gt, 1.0, 40.0), float(s[d > 0].mean()), (d, s))
    print(f"fp32 acc {res['fp32'][0]:.4f} score {res['fp32'][1]:.4f} | "
          f"{prec} acc {res[prec][0]:.4f} score {res[prec][1]:.4f}", flush=True)
    assert res[prec][0] >= res["fp32"][0] - 0.005
    assert res[prec][1] >= res["fp32"][1] - 0.005
    b = DepthmapBatch()
    for v in wide_scene.views:
        b.add_view(v.K, v.R, v.t, v.image, v.mask)
    b.add_job(0, list(range(1, 11)), wide_scene.min_depth, wide_scene.max_depth)
    b.set_patchmatch_iterations(2)
    b.set_min_patch_sd(1.0)
    b.set_seed(4)
    b.set_precision(prec)
    b.set_pyramid((2, 1))
    b.run()
    d, _, s, _ = b.result(0)
    assert np.array_equal(d, res[prec][2][0]) and np.array_equal(s, res[prec][2][1])


def test_precision_needs_patch_5_or_7(cfg1_scene):
    e = _add(DepthmapEstimator(), cfg1_scene, patch=9)
    e.set_precision("fp16")
    with pytest.raises(ValueError, match="patch size 5 or 7"):
        e.compute_patch_match()
    with pytest.raises(ValueError):
        e.set_precision(3)


def test_batch_precision_equals_single_view(wide_scene):
    print("batched FP16 vs single-view FP16", flush=True)
    b = DepthmapBatch()
    for v in wide_scene.views:
        b.add_view(v.K, v.R, v.t, v.image, v.mask)
    jobs = [(0, list(range(1, 11))), (5, [3, 4, 6, 7])]
    for ref, nb in jobs:
        b.add_job(ref, nb, wide_scene.min_depth, wide_scene.max_depth)
    b.set_patchmatch_iterations(2)
    b.set_min_patch_sd(1.0)
    b.set_seed(9)
    b.set_precision("fp16")
    b.run()
    for k, (ref, nb) in enumerate(jobs):
        e = DepthmapEstimator()
        for v in [ref] + nb:
            V = wide_scene.views[v]
            e.add_view(V.K, V.R, V.t, V.image, V.mask)
        e.set_depth_range(wide_scene.min_depth, wide_scene.max_depth, 50)
        e.set_patchmatch_iterations(2)
        e.set_min_patch_sd(1.0)
        e.set_seed(9 + k)
        e.set_precision("fp16")
        want = e.compute_patch_match()
        got = b.result(k)
        assert all(np.array_equal(x, y) for x, y in zip(got, want)), f"job {k}"


@pytest.fixture(scope="module")
def config3_scene():
    # 16 reference views of 2048x1536 with 10 neighbours each (26 cameras),
    # rendered on the GPU
    return scenes.scene_config3(device="cuda")


def test_config3_scorer_error_vs_fp64(config3_scene):
    """The error statement at the mixed-precision configuration's size: one
    2048x1536 reference view and its 10 neighbours, 1e5 pairs, the bounds
    above."""
    print("scorer error vs FP64 on 1e5 pairs, 2048x1536 + 10 neighbours", flush=True)
    ref, nb = config3_scene.jobs[0]
    sc = scenes.Scene(views=[config3_scene.views[v] for v in [ref] + nb],
                      min_depth=config3_scene.min_depth, max_depth=config3_scene.max_depth)
    ij, planes, views = _pairs(sc, N_PAIRS, seed=13)
    o = _add(oracle.OracleEstimator(), sc)
    want = np.array([o.score_f64(int(a), int(b), p.astype(np.float64), int(v))
                     for (a, b), p, v in zip(ij, planes, views)])
    e = _add(DepthmapEstimator(), sc)
    for prec in ("fp32", "fp16", "bf16"):
        e.set_precision(prec)
        got = e.score_hypotheses(ij, planes, views, fast=True).astype(np.float64)
        ok = (want > -1.0) & (got > -1.0)
        err = np.abs(got - want)[ok]
        flips = int(((want > -1.0) != (got > -1.0)).sum())
        print(f"{prec}: max {err.max():.3e} mean {err.mean():.3e} over {ok.sum()} pairs, "
              f"-1 flips {flips}", flush=True)
        assert ok.sum() > 0.3 * N_PAIRS
        assert err.max() <= BOUNDS[prec], (prec, err.max())
        assert flips <= 0.002 * N_PAIRS


def test_config3_batched_precisions(config3_scene):
    """All 16 jobs in one batch at 5 iterations in FP32, FP16 and BF16
    samples: valid maps, repeatable runs, and FP16/BF16 accuracy and mean
    score within 0.005 of FP32's."""
    sc = config3_scene
    b = DepthmapBatch()
    for v in sc.views:
        b.add_view(v.K, v.R, v.t, v.image, v.mask)
    for ref, nb in sc.jobs:
        b.add_job(ref, nb, sc.min_depth, sc.max_depth)
    b.set_patchmatch_iterations(5)
    b.set_min_patch_sd(1.0)
    b.set_seed(3)
    res = {}
    for prec in ("fp32", "fp16", "bf16"):
        print(f"config 3 batch, {prec}", flush=True)
        b.set_precision(prec)
        b.run()
        first = [b.result(k)[0].copy() for k in range(len(sc.jobs))]
        b.run()
        accs, scores = [], []
        for k, (ref, _) in enumerate(sc.jobs):
            d, p, s, n = b.result(k)
            assert np.array_equal(d, first[k])
            live = d > 0
            assert np.all(np.isfinite(d)) and live.mean() > 0.8
            assert np.all(s[live] >= -1) and (s[live] > 1 + 1e-3).mean() < 2e-3
            assert np.all((n[live] >= 1) & (n[live] <= 10))
            accs.append(scenes.depth_accuracy(d, sc.views[ref].depth, sc.min_depth, sc.max_depth))
            scores.append(float(s[live].mean()))
        res[prec] = (float(np.mean(accs)), float(np.mean(scores)))
        print(f"{prec}: accuracy {res[prec][0]:.4f} mean score {res[prec][1]:.4f}", flush=True)
    assert res["fp32"][0] > 0.9
    for prec in ("fp16", "bf16"):
        assert res[prec][0] >= res["fp32"][0] - 0.005
        assert res[prec][1] >= res["fp32"][1] - 0.005
