"""The learned PatchMatch initialiser's network, on the engine's own kernels.

A 4-level U-Net (channels 16, 32, 64, 128; 3x3 convolutions with leaky ReLU,
two per level at full and 1/2 resolution, three at 1/4, eight at the 1/8
bottleneck, 1.45 M parameters; max-pool down, nearest upsampling and skip
concatenation up; a 3x3 head with 4 outputs) whose every forward and backward step is a BF16
MFMA kernel of the engine (``pm_nn_*`` in include/opensfm_amd.h).  PyTorch
only holds the device memory and supplies the stream: no torch operator
computes a layer or a gradient.  Training uses the engine's Adam kernel.

Input: NHWC BF16, 40 channels (normalised image, a 32-plane NCC cost volume,
7 zero channels).  Output: FP32 (t, sx, sy, c) per pixel -- normalised log
depth, the slopes of the normal (sx, sy, -1) and a confidence logit.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _capi

IN_CHANNELS = 40
OUT_CHANNELS = 8   # 4 used; a multiple of 8 keeps the head's input gradient aligned
CHANNELS = (16, 32, 64, 128)
ENC = (2, 2, 3)    # convolutions per encoder level
DEC = (2, 2, 3)    # convolutions per decoder level (after the concatenation)
BOTTLENECK = 8     # convolutions at 1/8 resolution


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Ops:
    """Thin typed calls of the pm_nn_* kernels on torch device tensors."""

    def __init__(self):
        self.lib = _capi.load()

    def kpad(self, cin):
        return self.lib.pm_nn_conv_kpad(int(cin))

    def conv(self, x, x_f32, N, H, W, cin, xs, w16, cout, bias=None, y=None, y_f32=False,
             ys=0, yo=0, accumulate=False, act=None, as_=0, ao=0):
        _capi.check(self.lib.pm_nn_conv3x3(
            _p(x), int(x_f32), N * H * W, H, W, cin, xs, _p(w16), cout, self.kpad(cin),
            _p(bias), _p(y), int(y_f32), ys, yo, int(accumulate), _p(act), as_, ao, _stream()))

    def wgrad(self, x, N, H, W, cin, xs, dz, cout, dzs, dw, db):
        _capi.check(self.lib.pm_nn_conv3x3_wgrad(
            _p(x), N * H * W, H, W, cin, xs, _p(dz), cout, dzs, _p(dw), self.kpad(cin), _p(db),
            _stream()))

    def flip(self, w16, cout, cin, wt16):
        _capi.check(self.lib.pm_nn_weight_flip(_p(w16), cout, cin, self.kpad(cin), _p(wt16),
                                               self.kpad(cout), _stream()))

    def lrelu_backward(self, da, das, dao, a, as_, ao, dz, npix, c):
        _capi.check(self.lib.pm_nn_lrelu_backward(_p(da), das, dao, _p(a), as_, ao, _p(dz),
                                                  npix, c, _stream()))

    def maxpool(self, x, xs, xo, N, H, W, c, y, idx):
        _capi.check(self.lib.pm_nn_maxpool2(_p(x), xs, xo, N, H, W, c, _p(y), _p(idx),
                                            _stream()))

    def maxpool_backward(self, dy, idx, N, H, W, c, dx, dxs, dxo):
        _capi.check(self.lib.pm_nn_maxpool2_backward(_p(dy), _p(idx), N, H, W, c, _p(dx), dxs,
                                                     dxo, _stream()))

    def upsample(self, x, N, h, w, c, y, ys, yo):
        _capi.check(self.lib.pm_nn_upsample2(_p(x), N, h, w, c, _p(y), ys, yo, _stream()))

    def upsample_backward(self, dy, dys, dyo, N, h, w, c, dx):
        _capi.check(self.lib.pm_nn_upsample2_backward(_p(dy), dys, dyo, N, h, w, c, _p(dx),
                                                      _stream()))

    def loss(self, out, os, target, npix, n_valid, w_normal, w_conf, dout, sums):
        _capi.check(self.lib.pm_nn_loss(_p(out), os, _p(target), npix, int(n_valid),
                                        float(w_normal), float(w_conf), _p(dout), _p(sums),
                                        _stream()))

    def adam(self, w, w16, m, v, g, lr, step, b1=0.9, b2=0.999, eps=1e-8):
        _capi.check(self.lib.pm_nn_adam(_p(w), _p(w16), _p(m), _p(v), _p(g), w.numel(),
                                        float(lr), b1, b2, eps, int(step), _stream()))

    def to_bf16(self, x, y):
        _capi.check(self.lib.pm_nn_to_bf16(_p(x), _p(y), x.numel(), _stream()))


class Conv:
    """One 3x3 convolution: FP32 master weights [cout][kp] and bias, their
    BF16 copy, Adam moments and gradient accumulators."""

    def __init__(self, ops, cin, cout, device, rng, act=True):
        self.ops, self.cin, self.cout, self.act = ops, cin, cout, act
        kp = ops.kpad(cin)
        w = np.zeros((cout, kp), np.float32)
        std = np.sqrt(2.0 / (9 * cin)) if act else np.sqrt(1.0 / (9 * cin))
        w[:, :9 * cin] = rng.normal(0, std, (cout, 9 * cin))
        self.w = torch.from_numpy(w).to(device)
        self.b = torch.zeros(cout, dtype=torch.float32, device=device)
        self.w16 = torch.empty((cout, kp), dtype=torch.bfloat16, device=device)
        self.wt16 = torch.empty((cin, ops.kpad(cout)), dtype=torch.bfloat16, device=device)
        self.dw = torch.zeros_like(self.w)
        self.db = torch.zeros_like(self.b)
        self.mw, self.vw = torch.zeros_like(self.w), torch.zeros_like(self.w)
        self.mb, self.vb = torch.zeros_like(self.b), torch.zeros_like(self.b)
        self.refresh()

    def refresh(self):
        self.ops.to_bf16(self.w, self.w16)

    @property
    def n_params(self):
        return 9 * self.cin * self.cout + self.cout


class UNet:
    """See the module docstring.  forward() keeps what backward() needs.

    Layers: encoder level l (l = 0, 1, 2) has ENC[l] convolutions, the
    bottleneck (1/8 resolution, 128 channels) BOTTLENECK convolutions, and
    decoder level l DEC[l] convolutions after the skip concatenation; every
    convolution but the head is followed by a leaky ReLU."""

    def __init__(self, device="cuda", seed=0):
        self.ops = Ops()
        self.device = device
        rng = np.random.default_rng(seed)
        c = CHANNELS
        mk = lambda ci, co, act=True: Conv(self.ops, ci, co, device, rng, act)  # noqa: E731
        self.layers = {}
        cin = IN_CHANNELS
        for lvl in range(3):
            for k in range(ENC[lvl]):
                self.layers[f"e{lvl}_{k}"] = mk(cin if k == 0 else c[lvl], c[lvl])
            cin = c[lvl]
        for k in range(BOTTLENECK):
            self.layers[f"b_{k}"] = mk(c[2] if k == 0 else c[3], c[3])
        for lvl in (2, 1, 0):
            for k in range(DEC[lvl]):
                self.layers[f"d{lvl}_{k}"] = mk(c[lvl] + c[lvl + 1] if k == 0 else c[lvl], c[lvl])
        self.layers["head"] = mk(c[0], OUT_CHANNELS, act=False)
        self.step = 0

    @property
    def n_params(self):
        return sum(l.n_params for l in self.layers.values())

    # -- storage helpers
    def _bf(self, n, c):
        return torch.empty((n, c), dtype=torch.bfloat16, device=self.device)

    def _f(self, n, c, zero=False):
        f = torch.zeros if zero else torch.empty
        return f((n, c), dtype=torch.float32, device=self.device)

    def _chain_fwd(self, s, names, x, xs, N, H, W, npix, cout, last=None, last_s=0):
        """Runs the convolutions `names` on x (stride xs); the last writes
        into `last` (stride last_s) when given.  Returns (output, stride)."""
        for k, name in enumerate(names):
            L = self.layers[name]
            if k == len(names) - 1 and last is not None:
                out, os_ = last, last_s
            else:
                out, os_ = self._bf(npix, cout), cout
            self.ops.conv(x, False, N, H, W, L.cin, xs, L.w16, L.cout, bias=L.b, act=out,
                          as_=os_, ao=0)
            s[name] = (x, xs, out, os_)
            x, xs = out, os_
        return x, xs

    def forward(self, x, N, H, W):
        """x: BF16 (N*H*W, 40) NHWC; H, W multiples of 8.  Returns FP32
        (N*H*W, 8) outputs."""
        assert H % 8 == 0 and W % 8 == 0 and x.shape == (N * H * W, IN_CHANNELS)
        c = CHANNELS
        P = [N * H * W, N * H * W // 4, N * H * W // 16, N * H * W // 64]
        hs = [(H, W), (H // 2, W // 2), (H // 4, W // 4), (H // 8, W // 8)]
        s = {"N": N, "H": H, "W": W}
        cat = [self._bf(P[l], c[l] + c[l + 1]) for l in range(3)]
        s["cat"], s["idx"] = cat, []
        inp, ins = x, IN_CHANNELS
        for lvl in range(3):
            names = [f"e{lvl}_{k}" for k in range(ENC[lvl])]
            self._chain_fwd(s, names, inp, ins, N, *hs[lvl], P[lvl], c[lvl], cat[lvl],
                            c[lvl] + c[lvl + 1])
            p = self._bf(P[lvl + 1], c[lvl])
            idx = torch.empty((P[lvl + 1], c[lvl]), dtype=torch.uint8, device=self.device)
            self.ops.maxpool(cat[lvl], c[lvl] + c[lvl + 1], 0, N, *hs[lvl], c[lvl], p, idx)
            s["idx"].append(idx)
            inp, ins = p, c[lvl]
        low, lc = self._chain_fwd(s, [f"b_{k}" for k in range(BOTTLENECK)], inp, ins, N,
                                  *hs[3], P[3], c[3])
        for lvl in (2, 1, 0):
            cc = c[lvl] + c[lvl + 1]
            self.ops.upsample(low, N, *hs[lvl + 1], lc, cat[lvl], cc, c[lvl])
            low, lc = self._chain_fwd(s, [f"d{lvl}_{k}" for k in range(DEC[lvl])], cat[lvl], cc,
                                      N, *hs[lvl], P[lvl], c[lvl])
        out = self._f(P[0], OUT_CHANNELS)
        L = self.layers["head"]
        self.ops.conv(low, False, N, H, W, c[0], lc, L.w16, L.cout, bias=L.b, y=out, y_f32=True,
                      ys=OUT_CHANNELS, yo=0)
        s["head"] = (low, lc, out, OUT_CHANNELS)
        s["out"] = out
        self._saved = s
        return out

    def _conv_bwd(self, name, x, xs, N, H, W, dz, need_dx=True, dx=None, dxs=None):
        """Weight/bias gradients of layer `name` for output gradient dz
        (FP32, contiguous) and input x; returns the input gradient (FP32)."""
        L = self.layers[name]
        self.ops.wgrad(x, N, H, W, L.cin, xs, dz, L.cout, L.cout, L.dw, L.db)
        if not need_dx:
            return None
        self.ops.flip(L.w16, L.cout, L.cin, L.wt16)
        if dx is None:
            dx = self._f(N * H * W, L.cin)
            dxs = L.cin
        self.ops.conv(dz, True, N, H, W, L.cout, L.cout, L.wt16, L.cin, y=dx, y_f32=True,
                      ys=dxs, yo=0)
        return dx

    def _chain_bwd(self, names, d, ds, N, H, W, npix, need_first_dx=True):
        """Back through the activated convolutions `names` given the
        gradient d (stride ds) of the last one's output; returns the
        gradient of the chain's input (FP32, contiguous)."""
        s = self._saved
        for k in range(len(names) - 1, -1, -1):
            x, xs, a, as_ = s[names[k]]
            L = self.layers[names[k]]
            dz = self._f(npix, L.cout)
            self.ops.lrelu_backward(d, ds, 0, a, as_, 0, dz, npix, L.cout)
            d = self._conv_bwd(names[k], x, xs, N, H, W, dz, need_dx=k > 0 or need_first_dx)
            ds = L.cin
        return d

    def backward(self, dout):
        """Accumulates parameter gradients for dL/dout (FP32 (P0, 8))."""
        s = self._saved
        N, H, W = s["N"], s["H"], s["W"]
        c = CHANNELS
        P = [N * H * W, N * H * W // 4, N * H * W // 16, N * H * W // 64]
        hs = [(H, W), (H // 2, W // 2), (H // 4, W // 4), (H // 8, W // 8)]
        cat = s["cat"]
        x, xs, _, _ = s["head"]
        d = self._conv_bwd("head", x, xs, N, H, W, dout)
        dcat = [None, None, None]
        for lvl in (0, 1, 2):
            cc = c[lvl] + c[lvl + 1]
            dcat[lvl] = self._chain_bwd([f"d{lvl}_{k}" for k in range(DEC[lvl])], d, c[lvl], N,
                                        *hs[lvl], P[lvl])
            # gradient of the upsampled part -> the next-lower output
            d = self._f(P[lvl + 1], c[lvl + 1])
            self.ops.upsample_backward(dcat[lvl], cc, c[lvl], N, *hs[lvl + 1], c[lvl + 1], d)
        d = self._chain_bwd([f"b_{k}" for k in range(BOTTLENECK)], d, c[3], N, *hs[3], P[3])
        # encoder, deepest first: pool backward adds into the skip part of dcat
        for lvl in (2, 1, 0):
            cc = c[lvl] + c[lvl + 1]
            self.ops.maxpool_backward(d, s["idx"][lvl], N, *hs[lvl], c[lvl], dcat[lvl], cc, 0)
            d = self._chain_bwd([f"e{lvl}_{k}" for k in range(ENC[lvl])], dcat[lvl], cc, N,
                                *hs[lvl], P[lvl], need_first_dx=lvl > 0)

    def zero_grad(self):
        for L in self.layers.values():
            L.dw.zero_()
            L.db.zero_()

    def adam_step(self, lr):
        self.step += 1
        for L in self.layers.values():
            self.ops.adam(L.w, L.w16, L.mw, L.vw, L.dw, lr, self.step)
            self.ops.adam(L.b, None, L.mb, L.vb, L.db, lr, self.step)

    def state_dict(self):
        return {f"{k}.{n}": getattr(L, n).detach().cpu().numpy()
                for k, L in self.layers.items() for n in ("w", "b")}

    def load_state_dict(self, d):
        for k, L in self.layers.items():
            L.w.copy_(torch.from_numpy(np.asarray(d[f"{k}.w"], np.float32)))
            L.b.copy_(torch.from_numpy(np.asarray(d[f"{k}.b"], np.float32)))
            L.refresh()
