"""Pins of the initialiser network's CPU oracle (oracle/nn_oracle.py): each
layer against a direct-loop definition, and each backward against central
finite differences of the forward on tiny float64 tensors."""
import numpy as np

from oracle import nn_oracle as nn


def _fd(f, x, eps=1e-6):
    g = np.zeros_like(x)
    it = np.nditer(x, flags=["multi_index"])
    for _ in it:
        i = it.multi_index
        xp, xm = x.copy(), x.copy()
        xp[i] += eps
        xm[i] -= eps
        g[i] = (f(xp) - f(xm)) / (2 * eps)
    return g


def test_conv3x3_matches_direct_loops():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 5, 4, 3))
    w = rng.normal(size=(2, 9, 3))
    b = rng.normal(size=2)
    y = nn.conv3x3(x, w, b)
    N, H, W, _ = x.shape
    for n in range(N):
        for i in range(H):
            for j in range(W):
                for co in range(2):
                    s = b[co]
                    for tap in range(9):
                        ii, jj = i + tap // 3 - 1, j + tap % 3 - 1
                        if 0 <= ii < H and 0 <= jj < W:
                            s += x[n, ii, jj] @ w[co, tap]
                    assert abs(y[n, i, j, co] - s) < 1e-12


def test_conv3x3_gradients_by_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 4, 3, 2))
    w = rng.normal(size=(3, 9, 2))
    b = rng.normal(size=3)
    r = rng.normal(size=(1, 4, 3, 3))  # L = sum(r * conv(x))
    dx = nn.conv3x3_dgrad(r, w)
    dw, db = nn.conv3x3_wgrad(x, r)
    assert np.allclose(dx, _fd(lambda t: (r * nn.conv3x3(t, w, b)).sum(), x), atol=1e-6)
    assert np.allclose(dw, _fd(lambda t: (r * nn.conv3x3(x, t, b)).sum(), w), atol=1e-6)
    assert np.allclose(db, _fd(lambda t: (r * nn.conv3x3(x, w, t)).sum(), b), atol=1e-6)


def test_weight_flip_gives_the_input_gradient():
    rng = np.random.default_rng(2)
    w = rng.normal(size=(4, 9, 3))
    dy = rng.normal(size=(1, 5, 6, 4))
    assert np.allclose(nn.conv3x3(dy, nn.weight_flip(w)), nn.conv3x3_dgrad(dy, w), atol=1e-12)
    wp = nn.pack_weights(w)
    assert wp.shape == (4, nn.kpad(3)) and np.all(wp[:, 27:] == 0)
    assert np.array_equal(nn.unpack_weights(wp, 3), w)


def test_pool_upsample_lrelu_and_their_gradients():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(1, 4, 6, 2))
    y, idx = nn.maxpool2(x)
    for i in range(2):
        for j in range(3):
            blk = x[0, 2 * i:2 * i + 2, 2 * j:2 * j + 2]
            assert np.allclose(y[0, i, j], blk.reshape(4, 2).max(0))
    r = rng.normal(size=y.shape)
    assert np.allclose(nn.maxpool2_backward(r, idx, 4, 6),
                       _fd(lambda t: (r * nn.maxpool2(t)[0]).sum(), x), atol=1e-6)
    u = rng.normal(size=(1, 2, 3, 2))
    ru = rng.normal(size=(1, 4, 6, 2))
    assert np.allclose(nn.upsample2(u)[0, 3, 5], u[0, 1, 2])
    assert np.allclose(nn.upsample2_backward(ru), _fd(lambda t: (ru * nn.upsample2(t)).sum(), u),
                       atol=1e-6)
    z = rng.normal(size=(7,))
    a = nn.lrelu(z)
    assert np.allclose(nn.lrelu_backward(np.ones(7), a), np.where(z > 0, 1.0, 0.1))
    # first maximum wins on ties
    t = np.zeros((1, 2, 2, 1))
    assert nn.maxpool2(t)[1][0, 0, 0, 0] == 0


def test_loss_gradient_by_finite_differences():
    rng = np.random.default_rng(4)
    out = rng.normal(size=(9, 4)) * np.array([0.03, 1.0, 1.0, 2.0])
    tgt = np.concatenate([rng.normal(size=(9, 3)) * np.array([0.03, 1, 1]),
                          (rng.random((9, 1)) > 0.4).astype(float)], 1)
    tgt[0, 3] = 1.0
    terms, grad = nn.loss(out, tgt, 0.3, 0.2)
    fd = _fd(lambda t: sum(nn.loss(t, tgt, 0.3, 0.2)[0]), out, eps=1e-7)
    assert np.allclose(grad, fd, atol=1e-5)


def test_bf16_round():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.14159265], np.float32)
    r = nn.bf16_round(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties to even -> 1
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0078125)
    assert r[3] == -2.5 and abs(r[4] - 3.140625) < 1e-7
