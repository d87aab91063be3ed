"""The learned denoiser: a small U-Net trained on the GPU on pairs the renderer
itself produces (few-frame accumulations -> many-frame targets of seeded
procedural scenes), then attached to a Renderer so each frame's displayed image
is the network's output.

    python -m rustracer_amd.denoiser --scenes 24 --steps 600 --out weights.f32
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys
import time

import numpy as np

from . import RenderConfig, Renderer, Scene
from ._ffi import check, lib, ptr

INPUT_FRAMES = (1, 2, 4, 8, 16, 32)
# (cin, cout) of the six 3x3 convolutions: two at full resolution, two at half,
# two at full after the skip concatenation (the layout of csrc/unet.hip)
LAYER_SHAPES = ((16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32))


class UNet:
    def __init__(self, device: int = 0, seed: int = 1):
        h = C.c_void_p()
        check(lib().rt_unet_create(device, seed, C.byref(h)), "rt_unet_create")
        self._h = h

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
        check(lib().rt_unet_dataset_set(self._h, slot, which, renderer._h), "rt_unet_dataset_set")

    def train(self, steps: int, batch: int = 16, crop: int = 128, lr: float = 1e-3, seed: int = 1) -> np.ndarray:
        losses = np.zeros(max(steps, 1), np.float32)
        check(lib().rt_unet_train(self._h, steps, batch, crop, lr, seed, ptr(losses)), "rt_unet_train")
        return losses[:steps]

    def eval(self, features: np.ndarray, noisy: np.ndarray, target: np.ndarray | None = None):
        """One pass on host tensors (n, h, w, 16) bf16 bits / (n, h, w, 3) float32.
        Returns pred, and with a target also (grads, loss)."""
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


def _set(r, **fields) -> None:
    c = r.config
    for k, v in fields.items():
        setattr(c, k, v)
    r.config = c


def build_dataset(unet: UNet, n_scenes: int, size: int = 256, target_frames: int = 512, photons: int = 200_000,
                  first_seed: int = 1000, log=None) -> int:
    """Renders the training pairs into the network's dataset; returns the slot count."""
    unet.dataset_reset(size, size, n_scenes * len(INPUT_FRAMES))
    slot = 0
    for k in range(n_scenes):
        seed = first_seed + k
        scene = Scene.builtin("training", seed)
        cam = Scene.builtin_camera("training", seed, 0)
        cfg = RenderConfig(width=size, height=size, spp=1, max_depth=8, enable_denoise=False,
                           enable_photon_map=photons > 0, photon_count=max(photons, 1), accumulate_frames=1 << 20)
        with Renderer(cfg, scene, 0) as r:
            r.update_camera(cam)
            if photons > 0:
                r.rebuild_photons()
            r.render_frames(target_frames)
            first = slot
            for _ in INPUT_FRAMES:
                unet.dataset_set(slot, 1, r)  # the same target for every input of this scene
                slot += 1
            for j, frames in enumerate(INPUT_FRAMES):
                r.reset_accumulation()
                r.render_frames(frames)
                unet.dataset_set(first + j, 0, r)
        if log:
            log(f"training scene {k + 1}/{n_scenes} (seed {seed})")
    return slot


def train_denoiser(n_scenes: int = 24, steps: int = 600, batch: int = 16, crop: int = 128, lr: float = 1e-3,
                   seed: int = 1, log=None):
    """A trained UNet, the loss curve and the seconds spent (data, training)."""
    unet = UNet(0, seed)
    t0 = time.perf_counter()
    build_dataset(unet, n_scenes, log=log)
    t1 = time.perf_counter()
    losses = unet.train(steps, batch, crop, lr, seed)
    t2 = time.perf_counter()
    if log:
        log(f"trained {steps} steps in {t2 - t1:.1f} s (data {t1 - t0:.1f} s); "
            f"loss {losses[:10].mean():.4f} -> {losses[-10:].mean():.4f}")
    return unet, losses, (t1 - t0, t2 - t1)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=24)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)

    def log(m):
        print(f"[denoiser] {m}", file=sys.stderr, flush=True)

    unet, losses, _ = train_denoiser(args.scenes, args.steps, args.batch, args.crop, args.lr, log=log)
    if args.out:
        unet.get_params().tofile(args.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
