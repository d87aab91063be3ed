"""GPU: the learned denoiser's BF16 MFMA network against the CPU reference
(oracle/unet_ref.py, the same bf16 storage points, float32 arithmetic), and
training on renderer-produced pairs."""
import os
import sys

import numpy as np
import pytest

import rustracer_amd as rt
from rustracer_amd.denoiser import UNet, build_dataset

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import unet_ref as U  # noqa: E402

pytestmark = pytest.mark.gpu


def _params(seed):
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co in U.LAYERS:
        parts += [rng.normal(0, np.sqrt(2 / (9 * ci)), 9 * co * ci), rng.normal(0, 0.05, co)]
    parts += [rng.normal(0, 0.1, 3 * U.OUT_C), rng.normal(0, 0.05, 3)]
    return np.concatenate(parts).astype(np.float32)


def _inputs(seed, n, h, w):
    rng = np.random.default_rng(seed)
    f = rng.normal(0, 1, (n, h, w, U.IN_C)).astype(np.float32)
    f[..., 12:] = 0.0  # 12 feature channels, zero padding to 16
    x = rng.normal(0, 1, (n, h, w, 3)).astype(np.float32)
    t = (x + rng.normal(0, 0.3, x.shape)).astype(np.float32)
    return U.bf16_bits(f), x, t


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("n,h,w", [(2, 32, 64), (1, 48, 80)])
def test_forward_backward_match_reference(n, h, w):
    print(f"unet eval n={n} h={h} w={w}", flush=True)
    p = _params(1)
    fb, x, t = _inputs(2, n, h, w)
    net = UNet(0, 1)
    assert net.param_count() == U.param_count()
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    pred, grads, loss = net.eval(fb, x, t)
    rpred, rloss, cache = U.forward(p, U.from_bits(fb), x, t, quant=True)
    rgrad = U.backward(cache)
    # bf16 storage: GPU and reference may round a stored value differently when
    # their float32 sums (different orders) straddle a bf16 rounding boundary
    assert _rel(pred - x, rpred - x) < 1e-2
    assert abs(loss - rloss) < 1e-2 * rloss
    off = 0
    blocks = []
    for ci, co in U.LAYERS:
        blocks += [9 * co * ci, co]
    blocks += [3 * U.OUT_C, 3]
    for k, size in enumerate(blocks):
        a, b = grads[off:off + size], rgrad[off:off + size]
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))
        assert _rel(a, b) < 5e-2 and cos > 0.998, (k, _rel(a, b), cos)
        off += size
    net.close()


def test_zero_output_layer_returns_the_input_image():
    net = UNet(0, 3)  # a fresh network starts with a zero output layer
    fb, x, _ = _inputs(4, 1, 32, 32)
    pred = net.eval(fb, x)
    assert np.array_equal(pred, x)
    net.close()


def test_training_lowers_the_loss_and_the_held_out_error():
    """A short training run on renderer pairs lowers the training loss, and on
    a held-out scene the network's 2-frame image is closer to a 256-frame
    reference than the 2-frame accumulation itself."""
    net = UNet(0, 5)
    build_dataset(net, n_scenes=6, size=128, target_samples=256, photons=50_000, first_seed=300)
    losses = net.train(steps=200, batch=8, crop=96, lr=2e-3, seed=9)
    print(f"loss {losses[:10].mean():.4f} -> {losses[-10:].mean():.4f}", flush=True)
    assert losses[-20:].mean() < 0.8 * losses[:20].mean()
    scene = rt.Scene.builtin("training", 9999)
    cam = rt.Scene.builtin_camera("training", 9999, 0)
    cfg = rt.RenderConfig(width=128, height=128, spp=1, max_depth=8, enable_denoise=False,
                          enable_photon_map=True, photon_count=50_000, accumulate_frames=1)
    with rt.Renderer(cfg, scene, 0) as r:
        r.update_camera(cam)
        r.rebuild_photons()
        r.accum64_reset()
        for _ in range(256):
            r.record_frame()
            r.accum64_add()
        c = r.config
        c.accumulate_frames = 1 << 20
        r.config = c
        r.reset_accumulation()
        r.set_sample_count(5000)
        r.render_frames(2)
        plain = r.accum64_relmse_display(1e-2)
        r.set_denoiser(net)
        r.reset_accumulation()
        r.set_sample_count(5000)
        r.render_frames(2)
        learned = r.accum64_relmse_display(1e-2)
        disp = r.read_display()
        r.set_denoiser(None)
    print(f"held-out relMSE at 2 frames: plain {plain:.4f}, learned {learned:.4f}", flush=True)
    assert np.isfinite(disp).all()
    assert learned < plain
    net.close()


def test_chunked_frames_with_denoiser_match_frame_by_frame():
    """With the learned denoiser attached, frames rendered by rt_render_frames
    (eight per camera-pass launch, the network run after each frame) give the
    same network output, accumulation and rgba8 as the same frames issued one
    at a time with a sync in between."""
    net = UNet(0, 3)
    net.set_params(_params(7))
    scene = rt.Scene.builtin("cornell_box", 2)
    cam = rt.Scene.builtin_camera("cornell_box", 2)
    print("chunked frames with the learned denoiser", flush=True)
    res = []
    for chunked in (True, False):
        cfg = rt.RenderConfig(width=96, height=72, spp=1, max_depth=6, enable_denoise=True, accumulate_frames=8)
        with rt.Renderer(cfg, scene, 0) as r:
            r.update_camera(cam)
            r.set_denoiser(net)
            if chunked:
                r.render_frames(9)
            else:
                for _ in range(9):
                    r.record_frame()
                    r.synchronize()
            res.append((r.read_display(), r.read_accumulation(), r.read_output_rgba8(), r.sample_count()))
            r.set_denoiser(None)
    net.close()
    for name, a, b in zip(("display", "accumulation", "rgba8"), res[0][:3], res[1][:3]):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), name
    assert res[0][3] == res[1][3] == 9


@pytest.mark.parametrize("n,h,w", [(1, 48, 80), (2, 32, 96)])
def test_inference_fused_staging_equals_training_forward(n, h, w):
    """The inference pass stages the pool and the upsample inside the
    convolutions; its prediction equals the training forward's (separate pool
    and upsample kernels) bit for bit, including ragged tiles."""
    print(f"unet fused forward n={n} h={h} w={w}", flush=True)
    p = _params(4)
    fb, x, t = _inputs(5, n, h, w)
    net = UNet(0, 1)
    net.set_params(p)
    pred_train, _, _ = net.eval(fb, x, t)
    pred_infer = net.eval(fb, x, None)
    assert np.array_equal(pred_train.view(np.uint32), pred_infer.view(np.uint32))


@pytest.mark.parametrize("n,h,w", [(1, 48, 80)])
def test_every_convolution_variant_gives_the_same_bits(n, h, w):
    """Each layer shape has several compiled variants (input-channel slice
    width, output-channel blocking, weights staged once); all sum in one
    order, so prediction, loss and every gradient are bit-identical whichever
    runs, and the autotuner's timing-based choice never changes a result."""
    print(f"unet variants n={n} h={h} w={w}", flush=True)
    p = _params(8)
    fb, x, t = _inputs(9, n, h, w)
    net = UNet(0, 1)
    net.set_params(p)
    net.set_variant_mode(0)
    base = net.eval(fb, x, t)
    base_inf = net.eval(fb, x)
    # mode 0 folds every pool and upsample into the convolutions' staging; the
    # training forward runs them as kernels of their own
    assert np.array_equal(base_inf.view(np.uint32), base[0].view(np.uint32))
    for mode in range(1, 13):  # autotuned, then up to the eleventh compiled variant of each shape
        net.set_variant_mode(mode)
        pred, grads, loss = net.eval(fb, x, t)
        assert np.array_equal(pred.view(np.uint32), base[0].view(np.uint32)), mode
        assert np.array_equal(grads.view(np.uint32), base[1].view(np.uint32)), mode
        assert loss == base[2], mode
        assert np.array_equal(net.eval(fb, x).view(np.uint32), base_inf.view(np.uint32)), mode
    assert "->" in net.conv_choices()  # mode 1 tuned and recorded its picks
    net.close()
