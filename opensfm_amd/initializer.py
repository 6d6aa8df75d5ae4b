"""Learned PatchMatch initialiser: a U-Net seeds depth and normal.

PatchMatch starts from random planes (RandomInitialization,
opensfm/src/lib/dense/src/depthmap.cc:237-256) and needs several
propagation/refinement sweeps to converge.  Here a small U-Net
(opensfm_amd.nn, the engine's own BF16 MFMA kernels) reads the reference
image and a 32-plane NCC plane-sweep cost volume at 1/4 resolution and
predicts, per coarse pixel, a normalised log depth, the normal's slopes and
a confidence.  PatchMatch then starts from those planes where the network is
confident (logit >= 0) and from random planes elsewhere, and runs fewer
iterations.  Plain PatchMatch stays both the fallback (unconfident pixels,
and views the network is not confident about) and the judge: the network is
trained on the engine's own PatchMatch depthmaps, and evaluated against
rendered ground truth beside plain PatchMatch.

``python -m opensfm_amd.initializer --train`` renders synthetic scenes,
computes their PatchMatch depthmaps on the GPU, trains the network and
writes its weights to opensfm_amd/data/initializer_unet.npz.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

from . import nn, scenes
from .pydense import DepthmapEstimator

SCALE = 4
NPLANES = 32
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                       "initializer_unet.npz")


def coarse_grid(width, height, scale=SCALE):
    """(ow, oh): the coarse view size rounded up to multiples of 8."""
    wl, hl = width // scale, height // scale
    return (wl + 7) // 8 * 8, (hl + 7) // 8 * 8


def targets_from_patchmatch(depth, plane, min_depth, max_depth, ow, oh, scale=SCALE):
    """Coarse training targets (t, sx, sy, valid) from a PatchMatch result:
    the full-resolution pixel nearest each coarse cell's centre."""
    H, W = depth.shape
    tgt = np.zeros((oh, ow, 4), np.float32)
    ys = np.minimum(np.arange(oh) * scale + scale // 2, H - 1)
    xs = np.minimum(np.arange(ow) * scale + scale // 2, W - 1)
    d = depth[np.ix_(ys, xs)]
    p = plane[np.ix_(ys, xs)]
    inside = (np.arange(oh)[:, None] < H // scale) & (np.arange(ow)[None, :] < W // scale)
    valid = inside & (d > 0) & (p[..., 2] != 0)
    lmin, lmax = np.log(min_depth), np.log(max_depth)
    with np.errstate(divide="ignore", invalid="ignore"):
        tgt[..., 0] = np.where(valid, (np.log(np.maximum(d, 1e-6)) - lmin) / (lmax - lmin), 0)
        tgt[..., 1] = np.where(valid, np.clip(-p[..., 0] / p[..., 2], -20, 20), 0)
        tgt[..., 2] = np.where(valid, np.clip(-p[..., 1] / p[..., 2], -20, 20), 0)
    tgt[..., 3] = valid
    return tgt.reshape(-1, 4)


class LearnedInitializer:
    """Wraps the network for use with a DepthmapEstimator."""

    def __init__(self, weights=None, device="cuda", seed=0):
        self.net = nn.UNet(device=device, seed=seed)
        self.device = device
        if weights is not None:
            d = np.load(weights) if isinstance(weights, str) else weights
            self.net.load_state_dict(d)
        self._keep = None

    def network_input(self, est: DepthmapEstimator, width, height):
        ow, oh = coarse_grid(width, height)
        x = torch.empty((oh * ow, nn.IN_CHANNELS), dtype=torch.bfloat16, device=self.device)
        torch.cuda.synchronize()
        est.initializer_input(x, ow, oh, SCALE, NPLANES)
        return x, ow, oh

    def predict(self, est: DepthmapEstimator, width, height):
        x, ow, oh = self.network_input(est, width, height)
        out = self.net.forward(x, 1, oh, ow)
        return out, ow, oh

    def predict_batch(self, batch, job, width, height):
        """The network's map for job `job` of a DepthmapBatch (its input from
        the batch's shared coarse views); returns (map, ow, oh)."""
        ow, oh = coarse_grid(width, height)
        x = torch.empty((oh * ow, nn.IN_CHANNELS), dtype=torch.bfloat16, device=self.device)
        torch.cuda.synchronize()
        batch.initializer_input(job, x, ow, oh, SCALE, NPLANES)
        return self.net.forward(x, 1, oh, ow), ow, oh

    def attach(self, est: DepthmapEstimator, width, height):
        """Runs the network and hands its map to the estimator's next
        compute; returns the fraction of confident coarse cells."""
        out, ow, oh = self.predict(est, width, height)
        torch.cuda.synchronize()
        self._keep = out
        est.set_init_map(out, ow, oh, nn.OUT_CHANNELS, SCALE)
        wl, hl = width // SCALE, height // SCALE
        conf = out[:, 3].reshape(oh, ow)[:hl, :wl]
        return float((conf >= 0).float().mean())

    def detach(self, est: DepthmapEstimator):
        est.set_init_map(None)
        self._keep = None


def make_estimator(views, min_depth, max_depth, patch=7, sd=1.0, seed=0):
    e = DepthmapEstimator()
    for K, R, t, img, mask in views:
        e.add_view(K, R, t, img, mask)
    e.set_depth_range(min_depth, max_depth, 50)
    e.set_patch_size(patch)
    e.set_min_patch_sd(sd)
    e.set_seed(seed)
    return e


def scene_views(scene, ref, nb):
    return [(scene.views[v].K, scene.views[v].R, scene.views[v].t, scene.views[v].image,
             scene.views[v].mask) for v in [ref] + list(nb)]


def training_scenes(n, width, height, seed, nadir_fraction=1 / 3, device=None):
    """Arc scenes of boxes on terrain or a ground plane (the single-view
    family) and nadir fractal-terrain scenes, as (views, min, max) jobs."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n):
        s = int(rng.integers(1, 2 ** 31))
        if rng.random() < nadir_fraction:
            sc = scenes.nadir_scene(2, 3, width, height, seed=s, neighbours=5,
                                    altitude=float(rng.uniform(30, 50)), device=device)
            ref, nb = sc.jobs[int(rng.integers(0, len(sc.jobs)))]
            jobs.append((scene_views(sc, ref, nb), sc.min_depth, sc.max_depth, sc.views[ref]))
        else:
            terrain = (float(rng.uniform(0.5, 2.0)), float(rng.uniform(4, 10))) \
                if rng.random() < 0.6 else None
            sc = scenes.make_scene(width, height, int(rng.integers(4, 9)),
                                   int(rng.integers(4, 20)), s, 1.0, 40.0, terrain=terrain,
                                   elevation_deg=float(rng.uniform(25, 60)),
                                   spacing_deg=float(rng.uniform(5, 9)), device=device)
            jobs.append((scene_views(sc, 0, range(1, len(sc.views))), sc.min_depth,
                         sc.max_depth, sc.views[0]))
    return jobs


def build_samples(init, jobs, iterations=5):
    """Network inputs and PatchMatch targets for each job."""
    xs, ts = [], []
    for k, (views, dmin, dmax, ref) in enumerate(jobs):
        e = make_estimator(views, dmin, dmax, seed=1000 + k)
        e.set_patchmatch_iterations(iterations)
        depth, plane, _, _ = e.compute_patch_match()
        H, W = depth.shape
        x, ow, oh = init.network_input(e, W, H)
        xs.append(x)
        ts.append(torch.from_numpy(targets_from_patchmatch(depth, plane, dmin, dmax, ow, oh))
                  .to(init.device))
    return xs, ts, (ow, oh)


def train(init, xs, ts, grid, steps=1500, batch=8, lr=2e-3, w_normal=0.05, w_conf=0.2,
          log=None, seed=0):
    net = init.net
    ow, oh = grid
    rng = np.random.default_rng(seed)
    history = []
    for step in range(steps):
        sel = rng.choice(len(xs), size=min(batch, len(xs)), replace=False)
        x = torch.cat([xs[i] for i in sel])
        t = torch.cat([ts[i] for i in sel])
        nv = int(t[:, 3].sum().item())
        out = net.forward(x, len(sel), oh, ow)
        dout = torch.zeros_like(out)
        sums = torch.zeros(3, dtype=torch.float32, device=init.device)
        net.ops.loss(out, nn.OUT_CHANNELS, t, x.shape[0], nv, w_normal, w_conf, dout, sums)
        net.zero_grad()
        net.backward(dout)
        cur_lr = lr * (0.5 * (1 + np.cos(np.pi * step / steps)))
        net.adam_step(cur_lr)
        if step % 100 == 0 or step == steps - 1:
            terms = sums.cpu().numpy().tolist()
            history.append((step, terms))
            if log:
                log(f"step {step}: loss terms {np.round(terms, 5).tolist()}")
    return history


MODES = {"plain_5": (False, 5), "plain_2": (False, 2), "plain_1": (False, 1),
         "learned_1": (True, 1), "learned_2": (True, 2)}


def evaluate(init, jobs, modes=MODES):
    """Accuracy (fraction within 1 % of ground truth), mean score and wall
    time (network input, network and PatchMatch, device-synchronised) of
    PatchMatch from random planes and from the network's map, with the
    given iteration counts."""
    rows = []
    for k, (views, dmin, dmax, ref) in enumerate(jobs):
        r = {}
        H, W = ref.image.shape
        for mode, (learned, iters) in modes.items():
            e = make_estimator(views, dmin, dmax, seed=7000 + k)
            e.set_patchmatch_iterations(iters)
            if learned:  # coarse views and the network warmed outside the timed region
                init.attach(e, W, H)
                init.detach(e)
            e.run(1)  # warm
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            conf = init.attach(e, W, H) if learned else None
            t_net = time.perf_counter() - t0
            e.run(1)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            d, _, s, _ = e.download()
            if learned:
                init.detach(e)
            r[mode] = {"acc": scenes.depth_accuracy(d, ref.depth, dmin, dmax),
                       "score": float(s[d > 0].mean()) if (d > 0).any() else 0.0,
                       "ms": dt * 1e3, "mpix_s": W * H / dt / 1e6, "confident": conf,
                       "net_ms": t_net * 1e3}
        rows.append(r)
    out = {}
    for mode, (learned, _) in modes.items():
        out[mode] = {k: round(float(np.mean([r[mode][k] for r in rows])), 4)
                     for k in ("acc", "score", "ms", "mpix_s")}
        if learned:
            for k in ("confident", "net_ms"):
                out[mode][k] = round(float(np.mean([r[mode][k] for r in rows])), 4)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--train", action="store_true")
    ap.add_argument("--scenes", type=int, default=4000)
    ap.add_argument("--eval-scenes", type=int, default=12)
    ap.add_argument("--size", type=int, nargs=2, default=(320, 256))
    ap.add_argument("--eval-size", type=int, nargs=2, default=(1000, 752))
    ap.add_argument("--steps", type=int, default=12000)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--seed", type=int, default=123)
    ap.add_argument("--nadir-fraction", type=float, default=0.67)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--render-device", default="cuda",
                    help="device of the synthetic scene renderer (\"\" for the host)")
    args = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    W, H = args.size
    init = LearnedInitializer(seed=args.seed)
    if args.train:
        log(f"rendering {args.scenes} training scenes of {W}x{H}")
        t0 = time.perf_counter()
        xs, ts, grid = [], [], None
        # render and convert in chunks, so that only the network inputs and
        # targets of every scene stay resident
        for c0 in range(0, args.scenes, 200):
            jobs = training_scenes(min(200, args.scenes - c0), W, H, args.seed + c0,
                                   args.nadir_fraction, device=args.render_device or None)
            x, t, grid = build_samples(init, jobs)
            xs += x
            ts += t
            log(f"{len(xs)} scenes ({time.perf_counter() - t0:.1f} s)")
        log(f"training {init.net.n_params} parameters for {args.steps} steps")
        t0 = time.perf_counter()
        hist = train(init, xs, ts, grid, steps=args.steps, batch=args.batch, log=log)
        log(f"trained in {time.perf_counter() - t0:.1f} s")
        os.makedirs(os.path.dirname(args.out), exist_ok=True)
        np.savez(args.out, **init.net.state_dict())
        log(f"weights written to {args.out}")
    else:
        init = LearnedInitializer(weights=args.out)
        hist = None
    EW, EH = args.eval_size
    log(f"rendering {args.eval_scenes} nadir and {args.eval_scenes} arc evaluation views "
        f"of {EW}x{EH}")
    sc = scenes.nadir_scene(4, 4, EW, EH, seed=args.seed + 777, neighbours=10)
    nadir = [(scene_views(sc, r, nb), sc.min_depth, sc.max_depth, sc.views[r])
             for r, nb in sc.jobs[:args.eval_scenes]]
    arc = training_scenes(args.eval_scenes, EW, EH, args.seed + 999, nadir_fraction=0.0)
    res = {"nadir": evaluate(init, nadir), "arc": evaluate(init, arc)}
    print(json.dumps({"evaluation": res, "params": init.net.n_params,
                      "loss_history": hist}), flush=True)


if __name__ == "__main__":
    main()
