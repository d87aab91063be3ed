"""CPU reference of the learned denoiser's network (test infrastructure only).

The same U-Net as rustracer_amd/csrc/unet.hip, in numpy: 3x3 convolutions with
zero padding, ReLU, 2x2 average pool, nearest upsample, a 1x1 output layer with
a residual, and the mean squared log error. With `quant=True` every tensor the
GPU stores as bf16 (inputs of each convolution, activations, gradients between
layers) is rounded to bf16 (round to nearest even) at the same points, and the
arithmetic is float32; with `quant=False` everything is float64, which is the
version the tests check against finite differences.

Parameter vector layout (the GPU's): for each of the six 3x3 layers the weights
[9][cout][cin] (tap = 3*(dy+1) + (dx+1)) then the bias [cout]; then the output
layer's weights [3][32] and bias [3].
"""
from __future__ import annotations

import numpy as np

LAYERS = ((16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32))
OUT_C = 32


def param_count() -> int:
    return sum(9 * ci * co + co for ci, co in LAYERS) + 3 * OUT_C + 3


def unpack(p: np.ndarray):
    ws, bs, off = [], [], 0
    for ci, co in LAYERS:
        ws.append(p[off:off + 9 * co * ci].reshape(9, co, ci))
        off += 9 * co * ci
        bs.append(p[off:off + co])
        off += co
    ow = p[off:off + 3 * OUT_C].reshape(3, OUT_C)
    off += 3 * OUT_C
    ob = p[off:off + 3]
    return ws, bs, ow, ob


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bf16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def bf16_bits(x: np.ndarray) -> np.ndarray:
    return (bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def from_bits(b: np.ndarray) -> np.ndarray:
    return (b.astype(np.uint32) << 16).view(np.float32)


def _shift(x, dy, dx):
    """out[:, y, x] = x[:, y + dy, x + dx], zero outside."""
    n, h, w, c = x.shape
    p = np.zeros((n, h + 2, w + 2, c), x.dtype)
    p[:, 1:h + 1, 1:w + 1] = x
    return p[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w]


def conv(x, w):
    """y[p, co] = sum_tap sum_ci w[tap, co, ci] x[p + off(tap), ci]."""
    n, h, wd, _ = x.shape
    y = np.zeros((n, h, wd, w.shape[1]), x.dtype)
    for tap in range(9):
        y += _shift(x, tap // 3 - 1, tap % 3 - 1) @ w[tap].T
    return y


def conv_back_data(g, w):
    """dx[q, ci] = sum_tap sum_co w[tap, co, ci] g[q - off(tap), co]."""
    n, h, wd, _ = g.shape
    dx = np.zeros((n, h, wd, w.shape[2]), g.dtype)
    for tap in range(9):
        dx += _shift(g @ w[tap], 1 - tap // 3, 1 - tap % 3)
    return dx


def conv_back_weights(g, x):
    gw = np.zeros((9, g.shape[-1], x.shape[-1]), np.float64 if g.dtype == np.float64 else np.float32)
    g2 = g.reshape(-1, g.shape[-1])
    for tap in range(9):
        xs = _shift(x, tap // 3 - 1, tap % 3 - 1).reshape(-1, x.shape[-1])
        gw[tap] = g2.T @ xs
    return gw


def pool(x):
    s = ((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + x[:, 1::2, 0::2]) + x[:, 1::2, 1::2]
    return s * x.dtype.type(0.25)


def up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def forward(p, feats, noisy, target=None, quant=True):
    """feats (n, h, w, 16) values, noisy/target (n, h, w, 3) log values.
    Returns pred, loss (or None) and the cache for backward."""
    dt = np.float32 if quant else np.float64
    q = bf16 if quant else (lambda a: a)
    ws, bs, ow, ob = unpack(np.asarray(p, dt))
    if quant:
        ws = [bf16(w) for w in ws]
    relu = lambda a: np.maximum(a, 0)
    F = q(np.asarray(feats, dt))
    A1 = q(relu(conv(F, ws[0]) + bs[0]))
    A2 = q(relu(conv(A1, ws[1]) + bs[1]))
    P = q(pool(A2))
    B1 = q(relu(conv(P, ws[2]) + bs[2]))
    B2 = q(relu(conv(B1, ws[3]) + bs[3]))
    CAT = np.concatenate([A2, up(B2)], axis=-1)
    D1 = q(relu(conv(CAT, ws[4]) + bs[4]))
    D2 = q(relu(conv(D1, ws[5]) + bs[5]))
    pred = np.asarray(noisy, dt) + (D2 @ ow.T + ob)
    cache = dict(F=F, A1=A1, A2=A2, P=P, B1=B1, B2=B2, CAT=CAT, D1=D1, D2=D2, ws=ws, ow=ow, quant=quant)
    loss = None
    if target is not None:
        diff = pred - np.asarray(target, dt)
        loss = float(np.mean(diff.astype(np.float64) ** 2))
        cache["g"] = diff * dt(2.0 / diff.size)
    return pred, loss, cache


def backward(c):
    """Gradient of the loss w.r.t. the parameter vector (GPU order)."""
    quant = c["quant"]
    q = bf16 if quant else (lambda a: a)
    ws, ow = c["ws"], c["ow"]
    m = lambda a: (a > 0).astype(a.dtype)
    g = c["g"]
    gw = [None] * 6
    gb = [None] * 6
    gow = np.einsum("nhwc,nhwk->ck", g, c["D2"])
    gob = g.sum(axis=(0, 1, 2))
    gD2 = q(m(c["D2"]) * (g @ ow))

    def layer(k, G, X):
        gw[k] = conv_back_weights(G, X)
        gb[k] = G.sum(axis=(0, 1, 2))

    layer(5, gD2, c["D1"])
    gD1 = q(m(c["D1"]) * conv_back_data(gD2, ws[5]))
    layer(4, gD1, c["CAT"])
    gCAT = q(m(c["CAT"]) * conv_back_data(gD1, ws[4]))
    u = gCAT[..., 32:]
    gB2 = q(((u[:, 0::2, 0::2] + u[:, 0::2, 1::2]) + u[:, 1::2, 0::2]) + u[:, 1::2, 1::2])
    layer(3, gB2, c["B1"])
    gB1 = q(m(c["B1"]) * conv_back_data(gB2, ws[3]))
    layer(2, gB1, c["P"])
    gP = q(conv_back_data(gB1, ws[2]))
    gA2 = q(gCAT[..., :32] + up(gP) * gP.dtype.type(0.25) * m(c["A2"]))
    layer(1, gA2, c["A1"])
    gA1 = q(m(c["A1"]) * conv_back_data(gA2, ws[1]))
    layer(0, gA1, c["F"])
    parts = []
    for k in range(6):
        parts += [gw[k].ravel(), gb[k].ravel()]
    parts += [gow.ravel(), gob.ravel()]
    return np.concatenate(parts)
