"""The learned PatchMatch initialiser's network, on the engine's own kernels.

A 4-level U-Net (channels 16, 32, 64, 128; two 3x3 convolutions with leaky
ReLU per level; max-pool down, nearest upsampling and skip concatenation up;
a 3x3 head with 4 outputs) whose every forward and backward step is a BF16
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
    """See the module docstring.  forward() keeps what backward() needs."""

    def __init__(self, device="cuda", seed=0):
        self.ops = Ops()
        self.device = device
        rng = np.random.default_rng(seed)
        c0, c1, c2, c3 = CHANNELS
        mk = lambda ci, co, act=True: Conv(self.ops, ci, co, device, rng, act)  # noqa: E731
        self.layers = {
            "e0a": mk(IN_CHANNELS, c0), "e0b": mk(c0, c0),
            "e1a": mk(c0, c1), "e1b": mk(c1, c1),
            "e2a": mk(c1, c2), "e2b": mk(c2, c2),
            "ba": mk(c2, c3), "bb": mk(c3, c3),
            "d2a": mk(c2 + c3, c2), "d2b": mk(c2, c2),
            "d1a": mk(c1 + c2, c1), "d1b": mk(c1, c1),
            "d0a": mk(c0 + c1, c0), "d0b": mk(c0, c0),
            "head": mk(c0, OUT_CHANNELS, act=False),
        }
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

    def _conv_fwd(self, name, x, N, H, W, xs, act_out, as_, ao):
        L = self.layers[name]
        self.ops.conv(x, False, N, H, W, L.cin, xs, L.w16, L.cout, bias=L.b, act=act_out,
                      as_=as_, ao=ao)

    def forward(self, x, N, H, W):
        """x: BF16 (N*H*W, 40) NHWC; H, W multiples of 8.  Returns FP32
        (N*H*W, 8) outputs."""
        assert H % 8 == 0 and W % 8 == 0 and x.shape == (N * H * W, IN_CHANNELS)
        c0, c1, c2, c3 = CHANNELS
        s = {"N": N, "H": H, "W": W, "x": x}
        P = [N * H * W, N * H * W // 4, N * H * W // 16, N * H * W // 64]
        hs = [(H, W), (H // 2, W // 2), (H // 4, W // 4), (H // 8, W // 8)]
        cat = [self._bf(P[0], c0 + c1), self._bf(P[1], c1 + c2), self._bf(P[2], c2 + c3)]
        s["cat"] = cat
        # encoder
        a = self._bf(P[0], c0)
        self._conv_fwd("e0a", x, N, H, W, IN_CHANNELS, a, c0, 0)
        s["e0a"] = a
        self._conv_fwd("e0b", a, N, H, W, c0, cat[0], c0 + c1, 0)
        inp, ins = cat[0], c0 + c1
        pooled, idxs = [], []
        for lvl, (ca, cb, name) in enumerate([(c0, c1, "e1"), (c1, c2, "e2")]):
            h, w = hs[lvl]
            p = self._bf(P[lvl + 1], ca)
            idx = torch.empty((P[lvl + 1], ca), dtype=torch.uint8, device=self.device)
            self.ops.maxpool(inp, ins, 0, N, h, w, ca, p, idx)
            pooled.append(p)
            idxs.append(idx)
            h2, w2 = hs[lvl + 1]
            a = self._bf(P[lvl + 1], cb)
            self._conv_fwd(name + "a", p, N, h2, w2, ca, a, cb, 0)
            s[name + "a"] = a
            nxt = cat[lvl + 1]
            self._conv_fwd(name + "b", a, N, h2, w2, cb, nxt, nxt.shape[1], 0)
            inp, ins = nxt, nxt.shape[1]
        h, w = hs[2]
        p = self._bf(P[3], c2)
        idx = torch.empty((P[3], c2), dtype=torch.uint8, device=self.device)
        self.ops.maxpool(inp, ins, 0, N, h, w, c2, p, idx)
        pooled.append(p)
        idxs.append(idx)
        s["pooled"], s["idx"] = pooled, idxs
        h3, w3 = hs[3]
        ba = self._bf(P[3], c3)
        self._conv_fwd("ba", p, N, h3, w3, c2, ba, c3, 0)
        bb = self._bf(P[3], c3)
        self._conv_fwd("bb", ba, N, h3, w3, c3, bb, c3, 0)
        s["ba"], s["bb"] = ba, bb
        # decoder
        low, lc = bb, c3
        for lvl, (cs, cd, name) in [(2, (c2, c2, "d2")), (1, (c1, c1, "d1")), (0, (c0, c0, "d0"))]:
            h, w = hs[lvl]
            hl, wl = hs[lvl + 1]
            cc = cat[lvl].shape[1]
            self.ops.upsample(low, N, hl, wl, lc, cat[lvl], cc, cs)
            a = self._bf(P[lvl], cd)
            self._conv_fwd(name + "a", cat[lvl], N, h, w, cc, a, cd, 0)
            b = self._bf(P[lvl], cd)
            self._conv_fwd(name + "b", a, N, h, w, cd, b, cd, 0)
            s[name + "a"], s[name + "b"] = a, b
            low, lc = b, cd
        out = self._f(P[0], OUT_CHANNELS)
        L = self.layers["head"]
        self.ops.conv(low, False, N, H, W, c0, c0, L.w16, L.cout, bias=L.b, y=out, y_f32=True,
                      ys=OUT_CHANNELS, yo=0)
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

    def _act_bwd(self, da, das, dao, a, as_, ao, npix, c):
        dz = self._f(npix, c)
        self.ops.lrelu_backward(da, das, dao, a, as_, ao, dz, npix, c)
        return dz

    def backward(self, dout):
        """Accumulates parameter gradients for dL/dout (FP32 (P0, 8))."""
        s = self._saved
        N, H, W = s["N"], s["H"], s["W"]
        c0, c1, c2, c3 = CHANNELS
        P = [N * H * W, N * H * W // 4, N * H * W // 16, N * H * W // 64]
        hs = [(H, W), (H // 2, W // 2), (H // 4, W // 4), (H // 8, W // 8)]
        cat = s["cat"]
        # head (no activation)
        d = self._conv_bwd("head", s["d0b"], c0, N, H, W, dout)
        dcat = [None, None, None]
        for lvl, name, cd, cs in [(0, "d0", c0, c0), (1, "d1", c1, c1), (2, "d2", c2, c2)]:
            h, w = hs[lvl]
            cc = cat[lvl].shape[1]
            dz = self._act_bwd(d, cd, 0, s[name + "b"], cd, 0, P[lvl], cd)
            d = self._conv_bwd(name + "b", s[name + "a"], cd, N, h, w, dz)
            dz = self._act_bwd(d, cd, 0, s[name + "a"], cd, 0, P[lvl], cd)
            dcat[lvl] = self._conv_bwd(name + "a", cat[lvl], cc, N, h, w, dz)
            # gradient of the upsampled part -> the next-lower decoder output
            hl, wl = hs[lvl + 1]
            lc = c3 if lvl == 2 else (c2 if lvl == 1 else c1)
            d = self._f(P[lvl + 1], lc)
            self.ops.upsample_backward(dcat[lvl], cc, cs, N, hl, wl, lc, d)
        # d is now the gradient of bb (bottleneck output)
        h3, w3 = hs[3]
        dz = self._act_bwd(d, c3, 0, s["bb"], c3, 0, P[3], c3)
        d = self._conv_bwd("bb", s["ba"], c3, N, h3, w3, dz)
        dz = self._act_bwd(d, c3, 0, s["ba"], c3, 0, P[3], c3)
        d = self._conv_bwd("ba", s["pooled"][2], c2, N, h3, w3, dz)
        # encoder, deepest first: pool backward adds into the skip part of dcat
        for lvl, name, ca, cb in [(2, "e2", c1, c2), (1, "e1", c0, c1)]:
            h, w = hs[lvl]
            cc = cat[lvl].shape[1]
            self.ops.maxpool_backward(d, s["idx"][lvl], N, h, w, cb, dcat[lvl], cc, 0)
            dz = self._act_bwd(dcat[lvl], cc, 0, cat[lvl], cc, 0, P[lvl], cb)
            d = self._conv_bwd(name + "b", s[name + "a"], cb, N, h, w, dz)
            dz = self._act_bwd(d, cb, 0, s[name + "a"], cb, 0, P[lvl], cb)
            d = self._conv_bwd(name + "a", s["pooled"][lvl - 1], ca, N, h, w, dz)
        h, w = hs[0]
        cc = cat[0].shape[1]
        self.ops.maxpool_backward(d, s["idx"][0], N, h, w, c0, dcat[0], cc, 0)
        dz = self._act_bwd(dcat[0], cc, 0, cat[0], cc, 0, P[0], c0)
        d = self._conv_bwd("e0b", s["e0a"], c0, N, h, w, dz)
        dz = self._act_bwd(d, c0, 0, s["e0a"], c0, 0, P[0], c0)
        self._conv_bwd("e0a", s["x"], IN_CHANNELS, N, h, w, dz, need_dx=False)

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
