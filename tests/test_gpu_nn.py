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
    c = nn.CHANNELS
    hs = [(H, W), (H // 2, W // 2), (H // 4, W // 4), (H // 8, W // 8)]
    enc = [[f"e{l}_{k}" for k in range(nn.ENC[l])] for l in range(3)]
    dec = [[f"d{l}_{k}" for k in range(nn.DEC[l])] for l in range(3)]
    bot = [f"b_{k}" for k in range(nn.BOTTLENECK)]
    level = {n: l for l in range(3) for n in enc[l] + dec[l]}
    level.update({n: 3 for n in bot})
    t, inp = {}, {}

    def chain(names, v):
        for n in names:
            w, b = Ls[n]
            inp[n] = v
            v = r(O.lrelu(O.conv3x3(v, w, b)))
            t[n] = v
        return v

    v = x.reshape(N, H, W, nn.IN_CHANNELS)
    skips, idxs = [], []
    for l in range(3):
        v = chain(enc[l], v)
        skips.append(v)
        v, i = O.maxpool2(v)
        idxs.append(i)
    v = chain(bot, v)
    for l in (2, 1, 0):
        t[f"cat{l}"] = np.concatenate([skips[l], O.upsample2(v)], -1)
        v = chain(dec[l], t[f"cat{l}"])
    w, b = Ls["head"]
    inp["head"] = v
    out = O.conv3x3(v, w, b)
    if saved is not None:
        # the GPU's activations: every layer's input and output
        for n in list(inp):
            xg, xs, ag, as_ = saved[n]
            sh = hs[level.get(n, 0)]
            if n != "head":
                t[n] = npy(ag).reshape(N, *sh, as_)[..., :net.layers[n].cout]
            inp[n] = npy(xg).reshape(N, *sh, xs)[..., :net.layers[n].cin]
        idxs = [saved["idx"][l].cpu().numpy().reshape(N, *hs[l + 1], -1) for l in range(3)]

    def backward(dout):
        g = {}

        def layer(name, da, act=True):
            dz = O.lrelu_backward(da, t[name]) if act else da
            dw, db = O.conv3x3_wgrad(inp[name], r(dz))
            g[name] = (dw, O.conv3x3_wgrad(inp[name], dz)[1])
            return O.conv3x3_dgrad(r(dz), Ls[name][0])

        def chain_bwd(names, d):
            for n in reversed(names):
                d = layer(n, d)
            return d

        d = layer("head", dout.reshape(N, H, W, -1), act=False)
        dcat = [None] * 3
        for l in (0, 1, 2):
            dcat[l] = chain_bwd(dec[l], d)
            d = O.upsample2_backward(dcat[l][..., c[l]:])
        d = chain_bwd(bot, d)
        for l in (2, 1, 0):
            de = dcat[l][..., :c[l]] + O.maxpool2_backward(d, idxs[l], *hs[l])
            d = chain_bwd(enc[l], de)
        return g

    return out.reshape(N * H * W, -1), backward


def test_unet_forward_backward_matches_oracle():
    print("U-Net forward/backward vs oracle, 2 x 16 x 24", flush=True)
    net = nn.UNet(seed=3)
    assert 1.4e6 < net.n_params < 1.7e6
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
