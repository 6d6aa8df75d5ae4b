"""CPU oracle of the initialiser network's layers -- TEST INFRASTRUCTURE ONLY.

Plain numpy (float64) restatements of the operations the BF16 MFMA kernels in
opensfm_amd/csrc/nn_kernels.hip implement, forward and backward:
3x3 convolution (stride 1, zero padding 1) and its input / weight / bias
gradients, leaky ReLU (slope 0.1), 2x2 max pooling (first maximum wins),
nearest 2x upsampling, and the initialiser loss.  The learned initialiser
has no counterpart in the reference; these are the textbook definitions,
pinned in tests/test_nn_oracle.py by finite differences and by direct loops.

Tensors are NHWC.  Weights are (Cout, 9, Cin) with tap = 3 (dy + 1) + (dx + 1);
`pack_weights` gives the kernels' [Cout][Kp] layout (k = tap * Cin + ci).
"""
from __future__ import annotations

import numpy as np


def kpad(cin: int) -> int:
    return (9 * cin + 31) // 32 * 32


def bf16_round(x) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = rounded.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out).astype(np.float32)


def pack_weights(w: np.ndarray) -> np.ndarray:
    """(Cout, 9, Cin) -> (Cout, Kp) with k = tap * Cin + ci, zero padded."""
    cout, _, cin = w.shape
    out = np.zeros((cout, kpad(cin)), w.dtype)
    out[:, :9 * cin] = w.reshape(cout, 9 * cin)
    return out


def unpack_weights(wp: np.ndarray, cin: int) -> np.ndarray:
    return wp[:, :9 * cin].reshape(wp.shape[0], 9, cin)


def _shifted(x: np.ndarray, dy: int, dx: int) -> np.ndarray:
    """x[n, y + dy, x + dx, :] with zeros outside."""
    N, H, W, C = x.shape
    out = np.zeros_like(x)
    ys, yd = (slice(dy, H), slice(0, H - dy)) if dy >= 0 else (slice(0, H + dy), slice(-dy, H))
    xs, xd = (slice(dx, W), slice(0, W - dx)) if dx >= 0 else (slice(0, W + dx), slice(-dx, W))
    out[:, yd, xd, :] = x[:, ys, xs, :]
    return out


def conv3x3(x: np.ndarray, w: np.ndarray, b=None) -> np.ndarray:
    """y[n, i, j, co] = sum_{tap, ci} x[n, i + dy, j + dx, ci] w[co, tap, ci] + b[co]."""
    N, H, W, _ = x.shape
    y = np.zeros((N, H, W, w.shape[0]), np.float64)
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        y += _shifted(x.astype(np.float64), dy, dx) @ w[:, tap, :].T.astype(np.float64)
    if b is not None:
        y += np.asarray(b, np.float64)
    return y


def conv3x3_dgrad(dy: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Gradient of conv3x3 with respect to x."""
    dx = np.zeros(dy.shape[:3] + (w.shape[2],), np.float64)
    for tap in range(9):
        sy, sx = tap // 3 - 1, tap % 3 - 1
        dx += _shifted(dy.astype(np.float64), -sy, -sx) @ w[:, tap, :].astype(np.float64)
    return dx


def conv3x3_wgrad(x: np.ndarray, dy: np.ndarray):
    """Gradients of conv3x3 with respect to w (Cout, 9, Cin) and b."""
    cin, cout = x.shape[3], dy.shape[3]
    dw = np.zeros((cout, 9, cin), np.float64)
    d2 = dy.reshape(-1, cout).astype(np.float64)
    for tap in range(9):
        sy, sx = tap // 3 - 1, tap % 3 - 1
        xs = _shifted(x.astype(np.float64), sy, sx).reshape(-1, cin)
        dw[:, tap, :] = d2.T @ xs
    return dw, d2.sum(0)


def weight_flip(w: np.ndarray) -> np.ndarray:
    """The input-gradient convolution's weights: (Cin, 9, Cout) with
    wt[ci, tap, co] = w[co, 8 - tap, ci]."""
    return np.transpose(w[:, ::-1, :], (2, 1, 0)).copy()


def lrelu(z):
    return np.where(z > 0, z, 0.1 * z)


def lrelu_backward(da, a):
    return np.where(a > 0, da, 0.1 * da)


def maxpool2(x: np.ndarray):
    N, H, W, C = x.shape
    blocks = x.reshape(N, H // 2, 2, W // 2, 2, C).transpose(0, 1, 3, 5, 2, 4).reshape(
        N, H // 2, W // 2, C, 4)
    idx = np.argmax(blocks, axis=-1)  # first maximum
    return np.take_along_axis(blocks, idx[..., None], -1)[..., 0], idx


def maxpool2_backward(dy: np.ndarray, idx: np.ndarray, H: int, W: int) -> np.ndarray:
    N, h, w, C = dy.shape
    dx = np.zeros((N, H, W, C), np.float64)
    for s in range(4):
        m = idx == s
        dx[:, (s >> 1)::2, (s & 1)::2, :] += np.where(m, dy, 0.0)
    return dx


def upsample2(x: np.ndarray) -> np.ndarray:
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def upsample2_backward(dy: np.ndarray) -> np.ndarray:
    N, H, W, C = dy.shape
    return dy.reshape(N, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def _huber(d, delta):
    a = np.abs(d)
    val = np.where(a <= delta, 0.5 * d * d / delta, a - 0.5 * delta)
    grad = np.where(a <= delta, d / delta, np.sign(d))
    return val, grad


def loss(out: np.ndarray, target: np.ndarray, w_normal: float, w_conf: float):
    """out (P, >=4) = (t, sx, sy, c); target (P, 4) = (t*, sx*, sy*, valid).
    Returns the three loss terms and dL/dout (P, 4)."""
    out = out.astype(np.float64)
    valid = target[:, 3] > 0.5
    nv = max(1, int(valid.sum()))
    P = out.shape[0]
    l0, g0 = _huber(out[:, 0] - target[:, 0], 0.02)
    l1, g1 = _huber(out[:, 1] - target[:, 1], 0.5)
    l2, g2 = _huber(out[:, 2] - target[:, 2], 0.5)
    z = out[:, 3]
    y = valid.astype(np.float64)
    bce = np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z))) - y * z
    grad = np.zeros((P, 4))
    grad[:, 0] = np.where(valid, g0, 0) / nv
    grad[:, 1] = np.where(valid, w_normal * g1, 0) / nv
    grad[:, 2] = np.where(valid, w_normal * g2, 0) / nv
    grad[:, 3] = w_conf * (1 / (1 + np.exp(-z)) - y) / P
    terms = (float(np.where(valid, l0, 0).sum() / nv),
             float(w_normal * np.where(valid, l1 + l2, 0).sum() / nv),
             float(w_conf * bce.sum() / P))
    return terms, grad
