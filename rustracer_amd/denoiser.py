"""The learned denoiser: a five-level U-Net trained on the GPU on pairs the
renderer itself produces (1/2/4/8-frame accumulations -> 4096-sample targets of
seeded procedural scenes), then attached to a Renderer so each frame's
displayed image is the network's output.

    python -m rustracer_amd.denoiser --scenes 2000 --steps 2000 --out weights.f32
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import RenderConfig, Renderer, Scene, _ffi
from ._ffi import check, lib, ptr

INPUT_FRAMES = (1, 2, 4, 8)
TARGET_SAMPLES = 4096
TARGET_SPP = 16  # the target is rendered as TARGET_SAMPLES / TARGET_SPP frames of this many samples
# (cin, cout) of the 18 3x3 convolutions (the layout of csrc/unet.hip): two per
# encoder level (32, 64, 128, 128, 128 channels), then two per decoder level
# 3..0 reading [skip, upsampled] concatenations of 256, 256, 192, 96 channels
LAYER_SHAPES = ((16, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128), (128, 128), (128, 128),
                (128, 128), (128, 128), (256, 128), (128, 128), (256, 128), (128, 128), (192, 64), (64, 64),
                (96, 32), (32, 32))
# resolution level (0 = full, l = 1/2^l per axis) of each layer
LAYER_LEVELS = (0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 3, 3, 2, 2, 1, 1, 0, 0)
OUT_C = 32
PAD_TO = 16  # the network pads image sizes to multiples of this


class UNet:
    def __init__(self, device: int = 0, seed: int = 1):
        h = C.c_void_p()
        check(lib().rt_unet_create(device, seed, C.byref(h)), "rt_unet_create")
        self._h = h
        self._ds_lock = threading.Lock()

    def close(self) -> None:
        if self._h:
            lib().rt_unet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def param_count(self) -> int:
        n = C.c_size_t()
        check(lib().rt_unet_param_count(self._h, C.byref(n)), "rt_unet_param_count")
        return n.value

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.param_count(), np.float32)
        check(lib().rt_unet_get_params(self._h, ptr(out), out.size), "rt_unet_get_params")
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        check(lib().rt_unet_set_params(self._h, ptr(p), p.size), "rt_unet_set_params")

    def dataset_reset(self, width: int, height: int, capacity: int) -> None:
        check(lib().rt_unet_dataset_reset(self._h, width, height, capacity), "rt_unet_dataset_reset")

    def dataset_set(self, slot: int, which: int, renderer: Renderer) -> None:
        with self._ds_lock:  # one network, filled from several rendering threads
            check(lib().rt_unet_dataset_set(self._h, slot, which, renderer._h), "rt_unet_dataset_set")

    def train(self, steps: int, batch: int = 32, crop: int = 256, lr: float = 1e-3, seed: int = 1) -> np.ndarray:
        losses = np.zeros(max(steps, 1), np.float32)
        check(lib().rt_unet_train(self._h, steps, batch, crop, lr, seed, ptr(losses)), "rt_unet_train")
        return losses[:steps]

    def set_variant_mode(self, mode: int) -> None:
        """1 autotuned (default), 0 the first compiled variant per shape, k >= 2 the (k-2)-th."""
        check(lib().rt_unet_set_variant_mode(self._h, mode), "rt_unet_set_variant_mode")

    def conv_choices(self) -> str:
        """The convolution variant chosen per layer shape (autotuned at first use)."""
        n = C.c_size_t()
        check(lib().rt_unet_conv_choices(self._h, None, 0, C.byref(n)), "rt_unet_conv_choices")
        buf = C.create_string_buffer(n.value)
        check(lib().rt_unet_conv_choices(self._h, buf, n.value, None), "rt_unet_conv_choices")
        return buf.value.decode()

    def eval(self, features: np.ndarray, noisy: np.ndarray, target: np.ndarray | None = None):
        """One pass on host tensors (n, h, w, 16) bf16 bits / (n, h, w, 3) float32,
        h and w multiples of 16. Returns pred, and with a target also (grads, loss)."""
        n, h, w, _ = features.shape
        f = np.ascontiguousarray(features, np.uint16)
        x = np.ascontiguousarray(noisy, np.float32)
        t = None if target is None else np.ascontiguousarray(target, np.float32)
        pred = np.zeros((n, h, w, 3), np.float32)
        grads = np.zeros(self.param_count(), np.float32)
        loss = C.c_double()
        check(lib().rt_unet_eval(self._h, n, h, w, ptr(f), ptr(x), None if t is None else ptr(t), ptr(pred),
                                 ptr(grads), C.byref(loss)), "rt_unet_eval")
        return pred if t is None else (pred, grads, loss.value)


def network_flops(w: int, h: int) -> float:
    """Multiply-add FLOPs of one forward pass at w x h (padded to multiples of
    16): the 3x3 layers (12 of the first layer's 16 input channels carry data)
    and the 1x1 output layer."""
    pw, ph = -(-w // PAD_TO) * PAD_TO, -(-h // PAD_TO) * PAD_TO
    f = 0.0
    for k, ((cin, cout), lvl) in enumerate(zip(LAYER_SHAPES, LAYER_LEVELS)):
        cin_used = 12 if k == 0 else cin
        f += 2.0 * (pw >> lvl) * (ph >> lvl) * 9 * cin_used * cout
    return f + 2.0 * pw * ph * OUT_C * 3


def _render_scene_pair(unet: UNet, k: int, seed: int, size: int, target_samples: int, photons: int) -> None:
    """Scene `seed` into dataset slots k*4 .. k*4+3: the target (target_samples
    samples) and the 1/2/4/8-frame inputs (1 spp per frame)."""
    scene = Scene.builtin("training", seed)
    cam = Scene.builtin_camera("training", seed, 0)
    cfg = RenderConfig(width=size, height=size, spp=TARGET_SPP, max_depth=8, enable_denoise=False,
                       enable_photon_map=photons > 0, photon_count=max(photons, 1), accumulate_frames=1 << 20,
                       flags=_ffi.FLAG_MOMENT_HISTORY)
    first = k * len(INPUT_FRAMES)
    with Renderer(cfg, scene, 0) as r:
        r.update_camera(cam)
        if photons > 0:
            r.rebuild_photons()
        r.render_frames(max(1, target_samples // TARGET_SPP))
        for j in range(len(INPUT_FRAMES)):
            unet.dataset_set(first + j, 1, r)  # the same target for every input of this scene
        c = r.config
        c.spp = 1
        r.config = c
        for j, frames in enumerate(INPUT_FRAMES):
            r.reset_accumulation()
            r.render_frames(frames)
            unet.dataset_set(first + j, 0, r)


def build_dataset(unet: UNet, n_scenes: int, size: int = 256, target_samples: int = TARGET_SAMPLES,
                  photons: int = 200_000, first_seed: int = 1000, workers: int = 8, log=None) -> int:
    """Renders the training pairs of scenes first_seed .. first_seed + n_scenes - 1
    into the network's dataset (slot 4k + j holds scene k's input j); returns
    the slot count. Scenes are built and rendered by `workers` threads (the BVH
    build runs on the host); every slot's contents depend only on its scene."""
    unet.dataset_reset(size, size, n_scenes * len(INPUT_FRAMES))
    done = [0]
    lock = threading.Lock()

    def one(k):
        _render_scene_pair(unet, k, first_seed + k, size, target_samples, photons)
        if log:
            with lock:
                done[0] += 1
                if done[0] % max(1, n_scenes // 20) == 0 or done[0] == n_scenes:
                    log(f"training scenes {done[0]}/{n_scenes}")

    with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        for f in [ex.submit(one, k) for k in range(n_scenes)]:
            f.result()
    return n_scenes * len(INPUT_FRAMES)


def train_denoiser(n_scenes: int = 64, steps: int = 800, batch: int = 32, crop: int = 256, lr: float = 1e-3,
                   seed: int = 1, target_samples: int = TARGET_SAMPLES, log=None):
    """A trained UNet, the loss curve and the seconds spent (data, training)."""
    unet = UNet(0, seed)
    t0 = time.perf_counter()
    build_dataset(unet, n_scenes, target_samples=target_samples, log=log)
    t1 = time.perf_counter()
    losses = unet.train(steps, batch, crop, lr, seed)
    t2 = time.perf_counter()
    if log:
        log(f"trained {steps} steps in {t2 - t1:.1f} s (data {t1 - t0:.1f} s); "
            f"loss {losses[:10].mean():.4f} -> {losses[-10:].mean():.4f}")
    return unet, losses, (t1 - t0, t2 - t1)


def evaluate_held_out(unet: UNet, seeds, size: int = 256, ref_frames: int = 4096, photons: int = 200_000,
                      eps: float = 1e-2) -> dict:
    """relMSE (eps as in the convergence study) of the 1/2/4/8-frame
    accumulation and of the network's image against a ref_frames FP64
    reference, averaged over held-out training scenes `seeds`."""
    plain = {k: [] for k in INPUT_FRAMES}
    learned = {k: [] for k in INPUT_FRAMES}
    for seed in seeds:
        scene = Scene.builtin("training", seed)
        cam = Scene.builtin_camera("training", seed, 0)
        cfg = RenderConfig(width=size, height=size, spp=1, max_depth=8, enable_denoise=False,
                           enable_photon_map=photons > 0, photon_count=max(photons, 1), accumulate_frames=1)
        with Renderer(cfg, scene, 0) as r:
            r.update_camera(cam)
            if photons > 0:
                r.rebuild_photons()
            r.accum64_reset()
            for _ in range(ref_frames):
                r.record_frame()
                r.accum64_add()
            c = r.config
            c.accumulate_frames = 1 << 20
            r.config = c
            start = r.sample_count() + 1000
            for k in INPUT_FRAMES:
                for net, out in ((None, plain), (unet, learned)):
                    r.set_denoiser(net)
                    r.reset_accumulation()
                    r.set_sample_count(start)
                    r.render_frames(k)
                    out[k].append(r.accum64_relmse_display(eps))
            r.set_denoiser(None)
    return {"scenes": list(seeds), "reference_frames": ref_frames, "eps": eps,
            "relmse_accumulation": {k: float(np.mean(v)) for k, v in plain.items()},
            "relmse_network": {k: float(np.mean(v)) for k, v in learned.items()}}


def profile_passes(log, seed: int = 1) -> int:
    """Network passes on random inputs in a fixed order (for a kernel trace):
    autotune + 5 inference passes at 1920x1088, autotune + 3 training passes
    (forward + backward) at 32 x 256x256."""
    rng = np.random.default_rng(seed)
    unet = UNet(0, 1)

    def inputs(n, h, w, target):
        f = rng.normal(size=(n, h, w, 16)).astype(np.float32)
        f[..., 12:] = 0
        fb = (f.view(np.uint32) >> 16).astype(np.uint16)  # bf16 by truncation: any bits serve here
        x = rng.normal(size=(n, h, w, 3)).astype(np.float32)
        t = (x + 0.3 * rng.normal(size=x.shape)).astype(np.float32) if target else None
        return fb, x, t

    fb, x, _ = inputs(1, 1088, 1920, False)
    log("inference 1920x1088: autotune + 5 passes")
    for _ in range(6):
        unet.eval(fb, x)
    fb, x, t = inputs(32, 256, 256, True)
    log("training 32 x 256x256: autotune + 3 passes")
    for _ in range(4):
        unet.eval(fb, x, t)
    print(unet.conv_choices(), flush=True)
    unet.close()
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--crop", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--target-samples", type=int, default=TARGET_SAMPLES)
    ap.add_argument("--out", default=None, help="write the trained parameters (float32) here")
    ap.add_argument("--held-out", type=int, default=8, help="held-out scenes evaluated after training")
    ap.add_argument("--profile-passes", action="store_true",
                    help="only run network passes on random inputs, for a kernel trace: one autotuning pass and "
                         "five inference passes at 1920x1088, one autotuning and three training passes at "
                         "32 x 256x256; prints the chosen convolution variants")
    args = ap.parse_args(argv)

    def log(m):
        print(f"[denoiser] {m}", file=sys.stderr, flush=True)

    if args.profile_passes:
        return profile_passes(log)

    unet, losses, (data_s, train_s) = train_denoiser(args.scenes, args.steps, args.batch, args.crop, args.lr,
                                                     target_samples=args.target_samples, log=log)
    samples = args.scenes * 256 * 256 * (args.target_samples + sum(INPUT_FRAMES))
    res = {"scenes": args.scenes, "image": "256x256", "inputs": list(INPUT_FRAMES),
           "target_samples": args.target_samples, "dataset_seconds": round(data_s, 2),
           "dataset_gsamples_per_s": round(samples / data_s / 1e9, 3), "steps": args.steps,
           "batch": f"{args.batch} x {args.crop}x{args.crop}", "train_seconds": round(train_s, 2),
           "ms_per_step": round(train_s / max(args.steps, 1) * 1e3, 3),
           "loss_first10": float(losses[:10].mean()), "loss_last10": float(losses[-10:].mean()),
           "loss_curve": [float(losses[i:i + 100].mean()) for i in range(0, len(losses), max(1, len(losses) // 20))]}
    if args.out:
        unet.get_params().tofile(args.out)
    if args.held_out > 0:
        log(f"held-out evaluation on {args.held_out} scenes")
        res["held_out"] = evaluate_held_out(unet, range(900_000, 900_000 + args.held_out),
                                            ref_frames=args.target_samples)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
