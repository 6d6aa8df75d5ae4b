"""GPU tests of the initialiser network's BF16 MFMA kernels (pm_nn_*) against
the CPU oracle (oracle/nn_oracle.py), layer by layer and through the whole
U-Net's forward and backward.

The oracle runs in float64 on the same BF16-rounded operands the kernels
see (activations and weights are BF16; FP32 gradients are rounded to BF16
where they enter an MFMA).  Tolerances: BF16 outputs within 1 BF16 ulp
(relative 2^-7) plus 1e-3 absolute; FP32 outputs (gradients) within 1e-4
relative to the tensor's largest magnitude; the whole network's parameter
output within 2 % in norm; the backward, run by the oracle on the GPU
forward's own activations and pooling choices, within 1 % in norm per
parameter gradient.
"""
import numpy as np
import pytest
import torch

from oracle import nn_oracle as O
from opensfm_amd import nn

pytestmark = pytest.mark.gpu
DEV = "cuda"


def bf(a):
    return torch.from_numpy(O.bf16_round(a)).to(DEV).to(torch.bfloat16)


def f32(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(DEV)


def npy(t):
    return t.float().cpu().numpy().astype(np.float64)


@pytest.fixture(scope="module")
def ops():
    return nn.Ops()


@pytest.mark.parametrize("cin,cout,N,H,W", [(8, 16, 1, 5, 7), (16, 32, 2, 8, 8),
                                           (40, 16, 1, 12, 16), (48, 64, 1, 6, 10),
                                           (128, 8, 1, 4, 4)])
def test_conv_forward_and_gradients(ops, cin, cout, N, H, W):
    print(f"conv3x3 {cin}->{cout} on {N}x{H}x{W}", flush=True)
    rng = np.random.default_rng(cin * 100 + cout)
    x = O.bf16_round(rng.normal(size=(N, H, W, cin)))
    w = O.bf16_round(rng.normal(size=(cout, 9, cin)) * 0.1)
    b = rng.normal(size=cout).astype(np.float32)
    P = N * H * W
    # forward into a wider buffer at an offset, with the activation beside it
    ys, yo = cout + 8, 8
    y = torch.zeros((P, ys), dtype=torch.bfloat16, device=DEV)
    act = torch.zeros((P, cout), dtype=torch.bfloat16, device=DEV)
    xs = cin + 8
    xt = torch.zeros((P, xs), dtype=torch.bfloat16, device=DEV)
    xt[:, :cin] = bf(x.reshape(P, cin))
    w16 = bf(O.pack_weights(w))
    ops.conv(xt, False, N, H, W, cin, xs, w16, cout, bias=f32(b), y=y, ys=ys, yo=yo,
             act=act, as_=cout, ao=0)
    ref = O.conv3x3(x, w, b).reshape(P, cout)
    got = npy(y[:, yo:yo + cout])
    tol = np.abs(ref) * 2.0 ** -7 + 1e-3
    assert np.all(np.abs(got - ref) <= tol)
    assert np.all(np.abs(npy(act) - O.lrelu(ref)) <= np.abs(O.lrelu(ref)) * 2.0 ** -7 + 1e-3)
    # input gradient through the flipped weights, FP32 output
    dy = rng.normal(size=(N, H, W, cout)).astype(np.float32)
    wt16 = torch.empty((cin, ops.kpad(cout)), dtype=torch.bfloat16, device=DEV)
    ops.flip(w16, cout, cin, wt16)
    assert np.array_equal(npy(wt16), O.pack_weights(O.weight_flip(w)))
    dx = torch.zeros((P, cin), dtype=torch.float32, device=DEV)
    ops.conv(f32(dy.reshape(P, cout)), True, N, H, W, cout, cout, wt16, cin, y=dx, y_f32=True,
             ys=cin, yo=0)
    ref = O.conv3x3_dgrad(O.bf16_round(dy), w).reshape(P, cin)
    assert np.max(np.abs(npy(dx) - ref)) <= 1e-4 * np.max(np.abs(ref))
    # weight and bias gradients, accumulated onto existing values
    dw = torch.ones((cout, ops.kpad(cin)), dtype=torch.float32, device=DEV)
    db = torch.ones(cout, dtype=torch.float32, device=DEV)
    ops.wgrad(xt, N, H, W, cin, xs, f32(dy.reshape(P, cout)), cout, cout, dw, db)
    rw, rb = O.conv3x3_wgrad(x, O.bf16_round(dy))
    gw = O.unpack_weights(npy(dw), cin) - 1.0
    assert np.max(np.abs(gw - rw)) <= 1e-4 * np.max(np.abs(rw))
    assert np.all(npy(dw)[:, 9 * cin:] == 1.0)  # padding columns get no gradient
    rb_full = O.conv3x3_wgrad(x, dy.astype(np.float64))[1]
    assert np.max(np.abs(npy(db) - 1.0 - rb_full)) <= 1e-4 * np.max(np.abs(rb_full))


def test_pool_upsample_lrelu_loss_adam(ops):
    print("pool / upsample / lrelu / loss / adam", flush=True)
    rng = np.random.default_rng(7)
    N, H, W, C = 2, 6, 8, 16
    x = O.bf16_round(rng.normal(size=(N, H, W, C)))
    xs = C + 8
    xt = torch.zeros((N * H * W, xs), dtype=torch.bfloat16, device=DEV)
    xt[:, 8:] = bf(x.reshape(-1, C))
    y = torch.empty((N * H * W // 4, C), dtype=torch.bfloat16, device=DEV)
    idx = torch.empty((N * H * W // 4, C), dtype=torch.uint8, device=DEV)
    ops.maxpool(xt, xs, 8, N, H, W, C, y, idx)
    ry, ridx = O.maxpool2(x)
    assert np.array_equal(npy(y), ry.reshape(-1, C))
    assert np.array_equal(idx.cpu().numpy(), ridx.reshape(-1, C))
    dy = rng.normal(size=ry.shape).astype(np.float32)
    dx = torch.ones((N * H * W, xs), dtype=torch.float32, device=DEV)
    ops.maxpool_backward(f32(dy.reshape(-1, C)), idx, N, H, W, C, dx, xs, 8)
    assert np.allclose(npy(dx)[:, 8:] - 1.0,
                       O.maxpool2_backward(dy, ridx, H, W).reshape(-1, C), atol=1e-6)
    assert np.all(npy(dx)[:, :8] == 1.0)
    u = O.bf16_round(rng.normal(size=(N, H // 2, W // 2, C)))
    up = torch.zeros((N * H * W, xs), dtype=torch.bfloat16, device=DEV)
    ops.upsample(bf(u.reshape(-1, C)), N, H // 2, W // 2, C, up, xs, 8)
    assert np.array_equal(npy(up)[:, 8:], O.upsample2(u).reshape(-1, C))
    g = rng.normal(size=(N, H, W, C)).astype(np.float32)
    gt = torch.zeros((N * H * W, xs), dtype=torch.float32, device=DEV)
    gt[:, 8:] = f32(g.reshape(-1, C))
    du = torch.empty((N * H * W // 4, C), dtype=torch.float32, device=DEV)
    ops.upsample_backward(gt, xs, 8, N, H // 2, W // 2, C, du)
    assert np.allclose(npy(du), O.upsample2_backward(g).reshape(-1, C), atol=1e-5)
    dz = torch.empty((N * H * W, C), dtype=torch.float32, device=DEV)
    ops.lrelu_backward(gt, xs, 8, xt, xs, 8, dz, N * H * W, C)
    assert np.allclose(npy(dz), O.lrelu_backward(g, x).reshape(-1, C), atol=1e-7)
    # loss
    P = 1000
    out = rng.normal(size=(P, 8)).astype(np.float32) * np.array([0.05, 1, 1, 2, 0, 0, 0, 0],
                                                                  np.float32)
    tgt = np.concatenate([rng.normal(size=(P, 3)) * np.array([0.05, 1, 1]),
                          (rng.random((P, 1)) > 0.3)], 1).astype(np.float32)
    dout = torch.zeros((P, 8), dtype=torch.float32, device=DEV)
    sums = torch.zeros(3, dtype=torch.float32, device=DEV)
    nv = int((tgt[:, 3] > 0.5).sum())
    ops.loss(f32(out), 8, f32(tgt), P, nv, 0.3, 0.2, dout, sums)
    terms, grad = O.loss(out[:, :4], tgt, 0.3, 0.2)
    assert np.allclose(npy(sums), terms, rtol=1e-4)
    assert np.allclose(npy(dout)[:, :4], grad, rtol=1e-4, atol=1e-9)
    # adam against its definition
    w0 = rng.normal(size=100).astype(np.float32)
    gr = rng.normal(size=100).astype(np.float32)
    wt, m, v = f32(w0), torch.zeros(100, device=DEV), torch.zeros(100, device=DEV)
    w16 = torch.empty(100, dtype=torch.bfloat16, device=DEV)
    ops.adam(wt, w16, m, v, f32(gr), 0.01, 1)
    mh = 0.1 * gr / (1 - 0.9)
    vh = 0.001 * gr * gr / (1 - 0.999)
    assert np.allclose(npy(wt), w0 - 0.01 * mh / (np.sqrt(vh) + 1e-8), rtol=1e-5, atol=1e-6)
    assert np.array_equal(npy(w16), O.bf16_round(npy(wt)))


def _ref_unet(net, x, N, H, W, saved=None):
    """numpy forward/backward of the same U-Net with the kernels' BF16
    rounding points; returns the output and a function of dout giving the
    parameter gradients.  With `saved` (the GPU forward's tensors) the
    backward runs on the GPU's activations, pooling choices and
    concatenations, so that it checks the backward kernels alone."""
    Ls = {k: (O.unpack_weights(npy(L.w16), L.cin), npy(L.b)) for k, L in net.layers.items()}
    r = O.bf16_round
    c0, c1, c2, c3 = nn.CHANNELS

    def conv_act(name, t):
        w, b = Ls[name]
        return r(O.lrelu(O.conv3x3(t, w, b)))

    def sh(t, h, w):
        return t.reshape(N, h, w, -1)

    t = {}
    t["x"] = x.reshape(N, H, W, nn.IN_CHANNELS)
    t["e0a"] = conv_act("e0a", t["x"])
    t["e0b"] = conv_act("e0b", t["e0a"])
    t["p0"], t["i0"] = O.maxpool2(t["e0b"])
    t["e1a"] = conv_act("e1a", t["p0"])
    t["e1b"] = conv_act("e1b", t["e1a"])
    t["p1"], t["i1"] = O.maxpool2(t["e1b"])
    t["e2a"] = conv_act("e2a", t["p1"])
    t["e2b"] = conv_act("e2b", t["e2a"])
    t["p2"], t["i2"] = O.maxpool2(t["e2b"])
    t["ba"] = conv_act("ba", t["p2"])
    t["bb"] = conv_act("bb", t["ba"])
    t["cat2"] = np.concatenate([t["e2b"], O.upsample2(t["bb"])], -1)
    t["d2a"] = conv_act("d2a", t["cat2"])
    t["d2b"] = conv_act("d2b", t["d2a"])
    t["cat1"] = np.concatenate([t["e1b"], O.upsample2(t["d2b"])], -1)
    t["d1a"] = conv_act("d1a", t["cat1"])
    t["d1b"] = conv_act("d1b", t["d1a"])
    t["cat0"] = np.concatenate([t["e0b"], O.upsample2(t["d1b"])], -1)
    t["d0a"] = conv_act("d0a", t["cat0"])
    t["d0b"] = conv_act("d0b", t["d0a"])
    w, b = Ls["head"]
    out = O.conv3x3(t["d0b"], w, b)
    if saved is not None:
        hs = [(H, W), (H // 2, W // 2), (H // 4, W // 4), (H // 8, W // 8)]
        for k in ("e0a", "e1a", "e2a", "ba", "bb", "d2a", "d2b", "d1a", "d1b", "d0a", "d0b"):
            a = saved[k]
            lvl = {"e0a": 0, "e1a": 1, "e2a": 2, "ba": 3, "bb": 3, "d2a": 2, "d2b": 2,
                   "d1a": 1, "d1b": 1, "d0a": 0, "d0b": 0}[k]
            t[k] = npy(a).reshape(N, *hs[lvl], -1)
        for lvl, (cs, name) in enumerate([(c0, "e0b"), (c1, "e1b"), (c2, "e2b")]):
            cat = npy(saved["cat"][lvl]).reshape(N, *hs[lvl], -1)
            t[f"cat{lvl}"] = cat
            t[name] = cat[..., :cs]
        for lvl in range(3):
            t[f"p{lvl}"] = npy(saved["pooled"][lvl]).reshape(N, *hs[lvl + 1], -1)
            t[f"i{lvl}"] = saved["idx"][lvl].cpu().numpy().reshape(N, *hs[lvl + 1], -1)

    def backward(dout):
        g = {}

        def layer(name, xin, da, act=True):
            aout = t[name] if act else None
            dz = O.lrelu_backward(da, aout) if act else da
            dw, db = O.conv3x3_wgrad(xin, r(dz))
            g[name] = (dw, O.conv3x3_wgrad(xin, dz)[1])
            return O.conv3x3_dgrad(r(dz), Ls[name][0])

        d = layer("head", t["d0b"], dout.reshape(N, H, W, -1), act=False)
        d = layer("d0b", t["d0a"], d)
        dc0 = layer("d0a", t["cat0"], d)
        d = layer("d1b", t["d1a"], O.upsample2_backward(dc0[..., c0:]))
        dc1 = layer("d1a", t["cat1"], d)
        d = layer("d2b", t["d2a"], O.upsample2_backward(dc1[..., c1:]))
        dc2 = layer("d2a", t["cat2"], d)
        d = layer("bb", t["ba"], O.upsample2_backward(dc2[..., c2:]))
        d = layer("ba", t["p2"], d)
        de = dc2[..., :c2] + O.maxpool2_backward(d, t["i2"], H // 4, W // 4)
        d = layer("e2b", t["e2a"], de)
        d = layer("e2a", t["p1"], d)
        de = dc1[..., :c1] + O.maxpool2_backward(d, t["i1"], H // 2, W // 2)
        d = layer("e1b", t["e1a"], de)
        d = layer("e1a", t["p0"], d)
        de = dc0[..., :c0] + O.maxpool2_backward(d, t["i0"], H, W)
        d = layer("e0b", t["e0a"], de)
        layer("e0a", t["x"], d)
        return g

    return out.reshape(N * H * W, -1), backward


def test_unet_forward_backward_matches_oracle():
    print("U-Net forward/backward vs oracle, 2 x 16 x 24", flush=True)
    net = nn.UNet(seed=3)
    assert 4e5 < net.n_params < 2e6
    N, H, W = 2, 16, 24
    rng = np.random.default_rng(11)
    x = np.zeros((N * H * W, nn.IN_CHANNELS), np.float32)
    x[:, :33] = rng.normal(size=(N * H * W, 33))
    x = O.bf16_round(x)
    out = net.forward(bf(x), N, H, W)
    ref_out, _ = _ref_unet(net, x.astype(np.float64), N, H, W)
    got = npy(out)
    e_out = np.linalg.norm(got - ref_out) / np.linalg.norm(ref_out)
    print(f"  forward output rel err {e_out:.2e}", flush=True)
    assert e_out <= 0.02
    dout = rng.normal(size=(N * H * W, nn.OUT_CHANNELS)).astype(np.float32)
    net.zero_grad()
    net.backward(f32(dout))
    torch.cuda.synchronize()
    _, ref_bwd = _ref_unet(net, x.astype(np.float64), N, H, W, saved=net._saved)
    ref = ref_bwd(dout.astype(np.float64))
    worst = 0.0
    for name, L in net.layers.items():
        gw = O.unpack_weights(npy(L.dw), L.cin)
        rw, rb = ref[name]
        ew = np.linalg.norm(gw - rw) / np.linalg.norm(rw)
        eb = np.linalg.norm(npy(L.db) - rb) / max(1e-12, np.linalg.norm(rb))
        print(f"  {name}: weight grad rel err {ew:.2e}, bias {eb:.2e}", flush=True)
        worst = max(worst, ew, eb)
    assert worst <= 1e-2


def test_unet_training_reduces_loss():
    print("U-Net overfits a fixed batch", flush=True)
    net = nn.UNet(seed=5)
    N, H, W = 2, 16, 16
    rng = np.random.default_rng(2)
    x = np.zeros((N * H * W, nn.IN_CHANNELS), np.float32)
    x[:, :33] = rng.normal(size=(N * H * W, 33))
    tgt = np.zeros((N * H * W, 4), np.float32)
    tgt[:, 0] = 0.5 + 0.2 * np.tanh(x[:, 1])  # a function of the input
    tgt[:, 1] = np.tanh(x[:, 2])
    tgt[:, 3] = (x[:, 3] > 0).astype(np.float32)
    xt, tt = bf(x), f32(tgt)
    nv = int(tgt[:, 3].sum())
    losses = []
    for it in range(60):
        out = net.forward(xt, N, H, W)
        dout = torch.zeros_like(out)
        sums = torch.zeros(3, dtype=torch.float32, device=DEV)
        net.ops.loss(out, nn.OUT_CHANNELS, tt, N * H * W, nv, 0.3, 0.2, dout, sums)
        net.zero_grad()
        net.backward(dout)
        net.adam_step(3e-3)
        losses.append(float(sums.sum()))
    print(f"  loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.5 * losses[0]
