"""Pins of the CPU reference of the learned denoiser's network (oracle/unet_ref.py):
finite differences of the float64 version, adjoint identities of the
convolution kernels, and bf16 rounding known answers."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import unet_ref as U  # noqa: E402


def _random_problem(seed, n=1, h=16, w=16):
    rng = np.random.default_rng(seed)
    p = np.zeros(U.param_count())
    ws_off = 0
    for ci, co in U.LAYERS:
        p[ws_off:ws_off + 9 * co * ci] = rng.normal(0, np.sqrt(2 / (9 * ci)), 9 * co * ci)
        ws_off += 9 * co * ci
        p[ws_off:ws_off + co] = rng.normal(0, 0.1, co)
        ws_off += co
    p[ws_off:] = rng.normal(0, 0.2, p.size - ws_off)
    feats = rng.normal(0, 1, (n, h, w, U.IN_C))
    noisy = rng.normal(0, 1, (n, h, w, 3))
    target = rng.normal(0, 1, (n, h, w, 3))
    return p, feats, noisy, target


def test_backward_matches_finite_differences():
    p, f, x, t = _random_problem(3)
    _, loss, cache = U.forward(p, f, x, t, quant=False)
    g = U.backward(cache)
    assert g.shape == p.shape
    rng = np.random.default_rng(7)
    # a few parameters of every block: six conv weights/biases, output weights/bias
    offs, o = [], 0
    for ci, co in U.LAYERS:
        offs.append((o, 9 * co * ci))
        o += 9 * co * ci
        offs.append((o, co))
        o += co
    offs += [(o, 3 * U.OUT_C), (o + 3 * U.OUT_C, 3)]
    h = 1e-6
    for start, size in offs:
        for i in start + rng.choice(size, size=min(4, size), replace=False):
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (U.forward(pp, f, x, t, quant=False)[1] - U.forward(pm, f, x, t, quant=False)[1]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-8, (i, fd, g[i])


def test_conv_adjoint_identities():
    rng = np.random.default_rng(11)
    x = rng.normal(size=(2, 6, 10, 16))
    w = rng.normal(size=(9, 32, 16))
    y = rng.normal(size=(2, 6, 10, 32))
    lhs = np.sum(U.conv(x, w) * y)
    assert np.isclose(lhs, np.sum(x * U.conv_back_data(y, w)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(w * U.conv_back_weights(y, x)), rtol=1e-12)


def test_conv_of_a_delta_places_the_kernel():
    x = np.zeros((1, 5, 5, 16))
    x[0, 2, 2, 3] = 1.0
    w = np.arange(9 * 32 * 16, dtype=np.float64).reshape(9, 32, 16)
    y = U.conv(x, w)
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        # y[p] gathers x[p + off]: the delta at (2,2) reaches p = (2 - dy, 2 - dx)
        assert np.array_equal(y[0, 2 - dy, 2 - dx], w[tap, :, 3])


def test_bf16_rounding_known_answers():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, -2.5, 0.0], np.float32)
    got = U.bf16(v)
    assert got.tolist() == [1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -2.5, 0.0]
    assert np.array_equal(U.from_bits(U.bf16_bits(v)), got)


def test_untrained_output_layer_returns_its_input():
    p, f, x, _ = _random_problem(5)
    n_out = 3 * U.OUT_C + 3
    p[-n_out:] = 0.0
    pred, _, _ = U.forward(p, f, x, quant=True)
    assert np.array_equal(pred, x.astype(np.float32))
