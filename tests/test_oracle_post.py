"""Oracle a-trous (denoise.wgsl), composite (composite.wgsl) and tonemap
(tonemap.wgsl, examples/common/mod.rs:37-47) against closed forms. CPU only."""
import numpy as np
import pytest

from rustracer_amd import _ffi

from . import oracle_ffi as orc
from .helpers import alpha_for

F = np.float32
K1 = np.array([0.0625, 0.25, 0.375, 0.25, 0.0625])
RNG = np.random.default_rng(7)


def _params(w, h, step, alpha=1.0):
    return _ffi.DenoiseParams(w, h, step, 3.0, 12.0, 0.05, alpha, 0)


def _gbuf(h, w, normal=(0, 0, 1), depth=2.0):
    g = np.zeros((h, w, 4), F)
    g[..., :3] = normal
    g[..., 3] = depth
    return g


def _b3_reference(img, step):
    """Normalised B3 filter with the reference's in-bounds taps, float64."""
    h, w = img.shape[:2]
    out = np.zeros_like(img, dtype=np.float64)
    tot = np.zeros((h, w))
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            wgt = K1[dx + 2] * K1[dy + 2]
            ys, xs = np.mgrid[0:h, 0:w]
            ny, nx = ys + dy * step, xs + dx * step
            ok = (ny >= 0) & (nx >= 0) & (ny < h) & (nx < w)
            src = np.zeros_like(img, dtype=np.float64)
            src[ok] = img[ny[ok], nx[ok]]
            out += src * (wgt * ok)[..., None]
            tot += wgt * ok
    return out / tot[..., None]


@pytest.mark.parametrize("step", [1, 2, 4])
def test_atrous_is_b3_filter_when_unguided(step):
    h, w = 23, 31
    img = np.zeros((h, w, 4), F)
    img[..., :3] = RNG.random((h, w, 3), F)
    mom = np.zeros((h, w, 4), F)
    mom[..., 1] = 1e8  # huge variance: colour weight exp(-dl^2 / 3e8) == 1 to float precision
    out = orc.denoise_pass(img, _gbuf(h, w), mom, _params(w, h, step))
    ref = _b3_reference(img[..., :3].astype(np.float64), step)
    assert np.allclose(out[..., :3], ref, rtol=2e-6, atol=1e-7)
    assert (out[..., 3] == 1.0).all()


def test_atrous_constant_image_is_fixed_point():
    h, w = 16, 16
    img = np.ones((h, w, 4), F) * F(0.37)
    mom = RNG.random((h, w, 4), F)
    out = orc.denoise_pass(img, _gbuf(h, w), mom, _params(w, h, 2, 0.3))
    assert np.allclose(out[..., :3], 0.37, rtol=1e-6)


def test_atrous_zero_variance_keeps_detail():
    h, w = 20, 20
    img = np.zeros((h, w, 4), F)
    img[..., :3] = (RNG.integers(0, 64, (h, w, 1)) * 0.0125 + 0.1).astype(F)  # grey levels 0.0125 apart
    out = orc.denoise_pass(img, _gbuf(h, w), np.zeros((h, w, 4), F), _params(w, h, 1))
    # colour stop exp(-dl^2/1e-6): neighbours >= 0.0125 apart get weight < e^-150; equal ones agree
    assert np.allclose(out[..., :3], img[..., :3], rtol=1e-6)


def test_atrous_normal_and_depth_edges():
    h, w = 12, 24
    img = np.zeros((h, w, 4), F)
    img[:, :12, :3] = 0.2
    img[:, 12:, :3] = 0.9
    img[..., :3] += RNG.random((h, w, 3), F) * F(0.01)
    g = _gbuf(h, w)
    g[:, 12:, :3] = (1, 0, 0)  # orthogonal normals: pow(0, 12) = 0 across the edge
    mom = np.zeros((h, w, 4), F)
    mom[..., 1] = 1e8
    out = orc.denoise_pass(img, g, mom, _params(w, h, 1))
    left = _b3_reference(img[:, :12, :3].astype(np.float64), 1)
    assert np.allclose(out[:, :12, :3], left, rtol=2e-6)
    g2 = _gbuf(h, w)
    g2[:, 12:, 3] = 50.0  # depth jump: exp(-48/(0.05*2)) underflows to 0
    out2 = orc.denoise_pass(img, g2, mom, _params(w, h, 1))
    assert np.allclose(out2[:, :12, :3], left, rtol=2e-6)


def test_composite_running_mean_then_ema():
    h, w, N = 4, 5, 6
    frames = [RNG.random((h, w, 4), F) for _ in range(10)]
    acc = np.zeros((h, w, 4), F)
    for k, f in enumerate(frames, 1):
        acc = orc.composite(f, acc, alpha_for(N, k))
        if k <= N:  # alpha = 1/k: the arithmetic mean of the first k frames
            assert np.allclose(acc, np.mean(frames[:k], axis=0), rtol=1e-5)
    ema = np.mean(frames[:N], axis=0).astype(np.float64)
    for f in frames[N:]:
        ema = ema * (1 - 1 / N) + f * (1 / N)
    assert np.allclose(acc, ema, rtol=1e-5)
    assert alpha_for(0, 7) == 1.0


def test_tonemap_closed_form():
    h, w = 30, 40
    acc = np.zeros((h, w, 4), F)
    acc[..., :3] = np.exp(RNG.uniform(-6, 4, (h, w, 3))).astype(F)
    tm = _ffi.ToneMapUniform(1.3, 2.2, w, 1.0, (0, 0, 0, 0))
    out = orc.tonemap(acc, np.zeros((h, w, 4), F), tm)
    e = acc[..., :3].astype(np.float64) * 1.3
    v = np.clip((e / (e + 1)) ** (1 / 2.2), 0, 1) * 255
    ref = np.floor(v + 0.5)
    near = np.abs(v - np.floor(v) - 0.5) < 1e-3  # within float error of a rounding boundary
    d = out[..., :3].astype(int) - ref.astype(int)
    assert (d[~near] == 0).all() and (np.abs(d) <= 1).all()
    assert (out[..., 3] == 255).all()


def test_cpu_tonemap_truncates():
    acc = np.zeros((1, 6, 4), F)
    acc[0, :, 0] = [0.0, 0.5, 1.0, 4.0, 1e9, -1.0]
    out = orc.tonemap_cpu(acc)
    with np.errstate(divide="ignore", invalid="ignore"):
        v = np.clip(acc[0, :, 0].astype(np.float64) / (1 + acc[0, :, 0]), 0, 1) ** (1 / 2.2) * 255
    v = np.where(acc[0, :, 0] == -1.0, 0, v)  # -1 / 0 = -inf, clamped to 0
    assert list(out[0, :, 0]) == [int(x) for x in v]
    assert (out[..., 3] == 255).all()
