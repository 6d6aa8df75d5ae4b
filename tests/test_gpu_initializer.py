"""GPU tests of the learned initialiser's engine side and of what the
trained network gains.

* the network input (pm_estimator_initializer_input): channel 0 is the box-
  averaged reference image, channels 1..32 the coarse plane-sweep cost
  volume; checked against the oracle's FP64 ComputePlaneImageScore on the
  same coarse views (box-mean images, K scaled about pixel centres) within
  6e-3 (FP16 samples plus BF16 storage);
* a map built from ground truth seeds PatchMatch to high accuracy with no
  iteration at all, and a map with no confident cell reproduces plain
  PatchMatch bit for bit;
* the committed weights (opensfm_amd/data/initializer_unet.npz) gain over
  random initialisation at the same iteration count on held-out nadir views
  at the training resolution;
* the batch's network input equals the estimator's bit for bit, and the
  init map's arguments are checked;
* a batched job seeded with a map equals the single-view estimator seeded
  with the same map (and an unseeded job equals plain PatchMatch), also when
  the map seeds the coarsest level of a pyramid.
"""
import numpy as np
import pytest
import torch

import oracle
from opensfm_amd import DepthmapBatch, DepthmapEstimator, initializer as I, nn, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nadir():
    return scenes.nadir_scene(3, 3, 480, 352, seed=4242, neighbours=6)


def _est(sc, ref, nb, iters=1, seed=3):
    return I.make_estimator(I.scene_views(sc, ref, nb), sc.min_depth, sc.max_depth, seed=seed)


def _coarse(view, s=4):
    H, W = view.image.shape
    h, w = H // s, W // s
    blk = view.image[:h * s, :w * s].astype(np.int64).reshape(h, s, w, s).sum(axis=(1, 3))
    img = ((blk + s * s // 2) // (s * s)).astype(np.uint8)
    K = view.K.copy()
    K[:2, :] /= s
    K[0, 2] += (1.0 / s - 1.0) * 0.5
    K[1, 2] += (1.0 / s - 1.0) * 0.5
    return K, img


def test_network_input_matches_oracle(nadir):
    print("initialiser input vs oracle, 480x352 nadir + 6 neighbours", flush=True)
    ref, nb = nadir.jobs[4]
    e = _est(nadir, ref, nb)
    H, W = nadir.views[ref].image.shape
    ow, oh = I.coarse_grid(W, H)
    x = torch.empty((oh * ow, nn.IN_CHANNELS), dtype=torch.bfloat16, device="cuda")
    e.initializer_input(x, ow, oh, I.SCALE, I.NPLANES)
    got = x.float().cpu().numpy().reshape(oh, ow, nn.IN_CHANNELS)
    o = oracle.OracleEstimator()
    coarse = []
    for v in [ref] + list(nb):
        K, img = _coarse(nadir.views[v])
        o.add_view(K, nadir.views[v].R, nadir.views[v].t, img, np.full_like(img, 255))
        coarse.append(img)
    o.set_depth_range(nadir.min_depth, nadir.max_depth, 50)
    o.set_patch_size(7)
    h, w = coarse[0].shape
    assert np.array_equal(got[:h, :w, 0], (coarse[0].astype(np.float32) - 128) / 64)
    assert np.all(got[:, :, 33:] == 0) and np.all(got[h:, :, :] == 0) and np.all(got[:, w:, :] == 0)
    rng = np.random.default_rng(1)
    inv = np.linspace(1 / nadir.min_depth, 1 / nadir.max_depth, I.NPLANES)
    worst, n = 0.0, 0
    for _ in range(300):
        i, j = int(rng.integers(3, h - 3)), int(rng.integers(3, w - 3))
        d = int(rng.integers(0, I.NPLANES))
        plane = np.array([0.0, 0.0, -inv[d]])
        s = max(o.score_f64(i, j, plane, v) for v in range(1, len(nb) + 1))
        g = got[i, j, 1 + d]
        if s > -1 and g > -1:
            worst = max(worst, abs(g - s))
            n += 1
    print(f"  max |error| {worst:.2e} over {n} defined samples", flush=True)
    assert n > 200 and worst <= 6e-3


def test_init_map_from_ground_truth(nadir):
    print("init map from ground truth; all-unconfident map == plain", flush=True)
    ref, nb = nadir.jobs[4]
    V = nadir.views[ref]
    H, W = V.image.shape
    ow, oh = I.coarse_grid(W, H)
    # ground-truth map: depth at cell centres, normals from depth differences
    Kinv = np.linalg.inv(V.K)
    jj, ii = np.meshgrid(np.arange(W), np.arange(H))
    X = (np.stack([jj, ii, np.ones_like(jj)], -1) @ Kinv.T) * V.depth[..., None]
    dX = np.gradient(X, axis=1)
    dY = np.gradient(X, axis=0)
    N = np.cross(dY, dX)
    sx, sy = N[..., 0] / -N[..., 2], N[..., 1] / -N[..., 2]
    m = np.zeros((oh, ow, nn.OUT_CHANNELS), np.float32)
    ys = np.minimum(np.arange(oh) * 4 + 2, H - 1)
    xs = np.minimum(np.arange(ow) * 4 + 2, W - 1)
    d = V.depth[np.ix_(ys, xs)]
    lmin, lmax = np.log(nadir.min_depth), np.log(nadir.max_depth)
    m[..., 0] = (np.log(np.maximum(d, 1e-3)) - lmin) / (lmax - lmin)
    m[..., 1] = sx[np.ix_(ys, xs)]
    m[..., 2] = sy[np.ix_(ys, xs)]
    m[..., 3] = 10.0
    mt = torch.from_numpy(m.reshape(-1, nn.OUT_CHANNELS)).cuda()
    e = _est(nadir, ref, nb)
    e.set_patchmatch_iterations(0)
    e.set_init_spread(0.0, 0.0)
    e.set_init_map(mt, ow, oh, nn.OUT_CHANNELS, 4)
    acc0 = scenes.depth_accuracy(e.compute_patch_match()[0], V.depth, nadir.min_depth,
                                 nadir.max_depth)
    e.set_init_map(None)
    acc_rand = scenes.depth_accuracy(e.compute_patch_match()[0], V.depth, nadir.min_depth,
                                     nadir.max_depth)
    print(f"  0 iterations: GT-seeded {acc0:.4f}, random {acc_rand:.4f}", flush=True)
    assert acc0 > 0.8 and acc_rand < 0.1
    # no confident cell: same maps as plain PatchMatch
    m[..., 3] = -10.0
    mt = torch.from_numpy(m.reshape(-1, nn.OUT_CHANNELS)).cuda()
    e.set_patchmatch_iterations(2)
    plain = e.compute_patch_match()
    e.set_init_map(mt, ow, oh, nn.OUT_CHANNELS, 4)
    seeded = e.compute_patch_match()
    assert all(np.array_equal(a, b) for a, b in zip(plain, seeded))


@pytest.fixture(scope="module")
def nadir_train_scale():
    # held-out nadir views at the training resolution (1024x768)
    return scenes.nadir_scene(3, 3, 1024, 768, seed=4243, neighbours=6, device="cuda")


def test_trained_network_gains_over_random_init(nadir_train_scale):
    print("trained initialiser vs random initialisation, 1 iteration", flush=True)
    nadir = nadir_train_scale
    init = I.LearnedInitializer(weights=I.WEIGHTS)
    acc_l, acc_r = [], []
    for k, (ref, nb) in enumerate(nadir.jobs[:4]):
        V = nadir.views[ref]
        H, W = V.image.shape
        e = _est(nadir, ref, nb, seed=50 + k)
        e.set_patchmatch_iterations(1)
        d = e.compute_patch_match()[0]
        acc_r.append(scenes.depth_accuracy(d, V.depth, nadir.min_depth, nadir.max_depth))
        conf = init.attach(e, W, H)
        d = e.compute_patch_match()[0]
        init.detach(e)
        acc_l.append(scenes.depth_accuracy(d, V.depth, nadir.min_depth, nadir.max_depth))
        assert conf > 0.5
    print(f"  accuracy after 1 iteration: learned {np.mean(acc_l):.4f}, random {np.mean(acc_r):.4f}",
          flush=True)
    assert np.mean(acc_l) >= np.mean(acc_r) + 0.02


def test_batch_network_input_equals_estimator(nadir):
    """pm_batch_initializer_input (the pool's coarse views shared by every
    job) gives the estimator's network input bit for bit."""
    b = DepthmapBatch()
    for v in nadir.views:
        b.add_view(v.K, v.R, v.t, v.image, v.mask)
    for ref, nb in nadir.jobs[:3]:
        b.add_job(ref, nb, nadir.min_depth, nadir.max_depth)
    b.set_patch_size(7)
    for k, (ref, nb) in enumerate(nadir.jobs[:3]):
        H, W = nadir.views[ref].image.shape
        ow, oh = I.coarse_grid(W, H)
        xb = torch.empty((oh * ow, nn.IN_CHANNELS), dtype=torch.bfloat16, device="cuda")
        xe = torch.empty_like(xb)
        b.initializer_input(k, xb, ow, oh, I.SCALE, I.NPLANES)
        _est(nadir, ref, nb).initializer_input(xe, ow, oh, I.SCALE, I.NPLANES)
        assert torch.equal(xb.view(torch.int16), xe.view(torch.int16)), f"job {k}"


def test_init_map_arguments_are_checked(nadir):
    """set_init_map takes a float32 device tensor large enough for the grid
    and keeps it referenced (the engine reads it at the next compute)."""
    ref, nb = nadir.jobs[0]
    e = _est(nadir, ref, nb)
    with pytest.raises(TypeError):
        e.set_init_map(12345, 8, 8, nn.OUT_CHANNELS, I.SCALE)
    small = torch.zeros(10, dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        e.set_init_map(small, 8, 8, nn.OUT_CHANNELS, I.SCALE)
    with pytest.raises(ValueError):
        e.set_init_map(torch.zeros(8 * 8 * 8, dtype=torch.float32), 8, 8, nn.OUT_CHANNELS, I.SCALE)
    m = torch.zeros(8 * 8 * 8, dtype=torch.float32, device="cuda")
    e.set_init_map(m, 8, 8, nn.OUT_CHANNELS, I.SCALE)
    assert e._init_map is m
    e.set_init_map(None)
    assert e._init_map is None


def test_batched_learned_init_equals_estimator(nadir):
    print("batched learned initialisation vs single-view estimator", flush=True)
    init = I.LearnedInitializer(weights=I.WEIGHTS)
    seed = 90
    jobs = nadir.jobs[:3]
    b = DepthmapBatch()
    for v in nadir.views:
        b.add_view(v.K, v.R, v.t, v.image, v.mask)
    for ref, nb in jobs:
        b.add_job(ref, nb, nadir.min_depth, nadir.max_depth)
    b.set_patchmatch_iterations(1)
    b.set_patch_size(7)
    b.set_min_patch_sd(1.0)
    b.set_seed(seed)
    b.set_init_search(3)
    maps = {}
    for k in (0, 2):                      # job 1 stays unseeded
        ref, nb = jobs[k]
        H, W = nadir.views[ref].image.shape
        out, ow, oh = init.predict(_est(nadir, ref, nb, seed=seed + k), W, H)
        torch.cuda.synchronize()
        maps[k] = (out, ow, oh)
        b.set_init_map(k, out, ow, oh, nn.OUT_CHANNELS, I.SCALE)
    b.run()
    for k, (ref, nb) in enumerate(jobs):
        e = _est(nadir, ref, nb, seed=seed + k)
        e.set_patchmatch_iterations(1)
        e.set_init_search(3)
        if k in maps:
            out, ow, oh = maps[k]
            e.set_init_map(out, ow, oh, nn.OUT_CHANNELS, I.SCALE)
        want = e.compute_patch_match()
        got = b.result(k)
        assert all(np.array_equal(x, y) for x, y in zip(got, want)), f"job {k}"
    acc = [scenes.depth_accuracy(b.result(k)[0], nadir.views[r].depth, nadir.min_depth,
                                 nadir.max_depth) for k, (r, _) in enumerate(jobs)]
    print(f"  accuracy after 1 iteration: seeded {acc[0]:.4f} {acc[2]:.4f}, random {acc[1]:.4f}",
          flush=True)
    assert min(acc[0], acc[2]) > acc[1]
    with pytest.raises(IndexError):
        b.set_init_map(7, None)
    # with a pyramid the map seeds the coarsest level (scale 4 over two
    # coarser levels: one cell per pixel of the 1/4 level), in the batch and
    # in the estimator alike, bit for bit; 3 coarser levels need a scale
    # divisible by 8
    b.set_pyramid((1, 1))
    b.run()
    for k, (ref, nb) in enumerate(jobs):
        e = _est(nadir, ref, nb, seed=seed + k)
        e.set_patchmatch_iterations(1)
        e.set_init_search(3)
        e.set_pyramid((1, 1))
        if k in maps:
            out, ow, oh = maps[k]
            e.set_init_map(out, ow, oh, nn.OUT_CHANNELS, I.SCALE)
        want = e.compute_patch_match()
        got = b.result(k)
        assert all(np.array_equal(x, y) for x, y in zip(got, want)), f"pyramid job {k}"
        assert (got[0] > 0).mean() > 0.5
        if k in maps:
            e.set_pyramid((1, 1, 1))
            with pytest.raises(ValueError):
                e.compute_patch_match()
    b.set_pyramid((1, 1, 1))
    with pytest.raises(ValueError):
        b.run()
