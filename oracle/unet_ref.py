"""CPU reference of the learned denoiser's network (test infrastructure only).

The same U-Net as rustracer_amd/csrc/unet.hip, in numpy: five levels of 3x3
convolutions with zero padding and ReLU (32, 64, 128, 128, 128 channels), 2x2
average pools down, nearest upsamples concatenated after the skip tensor on the
way up (96, 192, 256, 256 input channels), a 1x1 output layer with a residual on
the log radiance, and the L1 loss. With `quant=True` every tensor the GPU stores
as bf16 (inputs of each convolution, activations, gradients between layers) is
rounded to bf16 (round to nearest even) at the same points, and the arithmetic
is float32; with `quant=False` everything is float64, which is the version the
tests check against finite differences.

Parameter vector layout (the GPU's): for each of the 18 3x3 layers the weights
[9][cout][cin] (tap = 3*(dy+1) + (dx+1)) then the bias [cout]; then the output
layer's weights [3][32] and bias [3].
"""
from __future__ import annotations

import numpy as np

IN_C = 16  # 12 feature channels, padded with zeros to the MFMA's K step
ENC = (32, 64, 128, 128, 128)  # channels per level, level 4 is the bottleneck
LEVELS = len(ENC)
# (cin, cout) of the 18 layers: two per encoder level, then two per decoder
# level 3..0 (the first of them reads [skip, upsampled] channels)
LAYERS = ((IN_C, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128), (128, 128), (128, 128),
          (128, 128), (128, 128),
          (256, 128), (128, 128), (256, 128), (128, 128), (192, 64), (64, 64), (96, 32), (32, 32))
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


def up_back(g):
    """Adjoint of `up`: the sum of each 2x2 block."""
    return ((g[:, 0::2, 0::2] + g[:, 0::2, 1::2]) + g[:, 1::2, 0::2]) + g[:, 1::2, 1::2]


def forward(p, feats, noisy, target=None, quant=True):
    """feats (n, h, w, 16) values, noisy/target (n, h, w, 3) log values; h and w
    multiples of 16. Returns pred, loss (or None) and the cache for backward."""
    dt = np.float32 if quant else np.float64
    q = bf16 if quant else (lambda a: a)
    ws, bs, ow, ob = unpack(np.asarray(p, dt))
    if quant:
        ws = [bf16(w) for w in ws]
    relu = lambda a: np.maximum(a, 0)
    layer = lambda k, x: q(relu(conv(x, ws[k]) + bs[k]))
    c = dict(ws=ws, ow=ow, quant=quant)
    c["F"] = x = q(np.asarray(feats, dt))
    skips = []
    for lvl in range(LEVELS):
        if lvl > 0:
            x = q(pool(skips[-1]))
            c[f"P{lvl}"] = x
        c[f"T{lvl}"] = t = layer(2 * lvl, x)
        x = layer(2 * lvl + 1, t)
        if lvl < LEVELS - 1:
            skips.append(x)
            c[f"S{lvl}"] = x
    c["B"] = x
    for lvl in range(LEVELS - 2, -1, -1):
        k = 2 * LEVELS + 2 * (LEVELS - 2 - lvl)
        cat = np.concatenate([skips[lvl], up(x)], axis=-1)
        c[f"CAT{lvl}"] = cat
        c[f"U{lvl}"] = u = layer(k, cat)
        c[f"V{lvl}"] = x = layer(k + 1, u)
    pred = np.asarray(noisy, dt) + (x @ ow.T + ob)
    loss = None
    if target is not None:
        diff = pred - np.asarray(target, dt)
        loss = float(np.mean(np.abs(diff.astype(np.float64))))
        c["g"] = np.sign(diff) * dt(1.0 / diff.size)
    return pred, loss, c


def backward(c):
    """Gradient of the L1 loss w.r.t. the parameter vector (GPU order)."""
    quant = c["quant"]
    q = bf16 if quant else (lambda a: a)
    ws, ow = c["ws"], c["ow"]
    m = lambda a: (a > 0).astype(a.dtype)
    g = c["g"]
    n_layers = len(LAYERS)
    gw, gb = [None] * n_layers, [None] * n_layers

    def wgrad(k, G, X):
        gw[k] = conv_back_weights(G, X)
        gb[k] = G.sum(axis=(0, 1, 2))

    V0 = c["V0"]
    gow = np.einsum("nhwc,nhwk->ck", g, V0)
    gob = g.sum(axis=(0, 1, 2))
    gx = q(m(V0) * (g @ ow))  # gradient at V_lvl (masked by its ReLU)
    # decoder, from the output (level 0) to the deepest concatenation; the
    # gradient of V_lvl (lvl > 0) is the 2x2 sum of the upsampled part of
    # CAT_{lvl-1}'s gradient, which already carries V_lvl's ReLU mask
    gcat = {}
    for lvl in range(0, LEVELS - 1):
        k = 2 * LEVELS + 2 * (LEVELS - 2 - lvl)
        if lvl > 0:
            gx = q(up_back(gcat[lvl - 1][..., ENC[lvl - 1]:]))
        wgrad(k + 1, gx, c[f"U{lvl}"])
        gu = q(m(c[f"U{lvl}"]) * conv_back_data(gx, ws[k + 1]))
        wgrad(k, gu, c[f"CAT{lvl}"])
        gcat[lvl] = q(m(c[f"CAT{lvl}"]) * conv_back_data(gu, ws[k]))
    # encoder, from the bottleneck B up: the gradient of a skip tensor S is its
    # part of the concatenation's gradient plus the pool's share, masked by S
    gx = q(up_back(gcat[LEVELS - 2][..., ENC[LEVELS - 2]:]))  # gradient at B
    for lvl in range(LEVELS - 1, -1, -1):
        t = c[f"T{lvl}"]
        wgrad(2 * lvl + 1, gx, t)
        gt = q(m(t) * conv_back_data(gx, ws[2 * lvl + 1]))
        xin = c["F"] if lvl == 0 else c[f"P{lvl}"]
        wgrad(2 * lvl, gt, xin)
        if lvl == 0:
            break
        gp = q(conv_back_data(gt, ws[2 * lvl]))
        s = c[f"S{lvl - 1}"]
        cs = ENC[lvl - 1]
        gx = q(gcat[lvl - 1][..., :cs] + up(gp) * gp.dtype.type(0.25) * m(s))
    parts = []
    for k in range(n_layers):
        parts += [gw[k].ravel(), gb[k].ravel()]
    parts += [gow.ravel(), gob.ravel()]
    return np.concatenate(parts)
