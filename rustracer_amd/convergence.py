"""Convergence study: wall time until the displayed image reaches relMSE <= 0.01.

relMSE = mean over pixels and rgb of (x - y)^2 / (y^2 + eps), eps = 1e-2, where y is
an FP64 reference (the mean of `ref_frames` independent 1-spp frames, accumulated
in double on the GPU) and x is the displayed linear image (the accumulation
buffer) of the method under test. A method is run once frame by frame with the
relMSE taken after every frame (untimed) to find the first frame count k that
reaches the threshold, then the same k frames are re-run back to back from the
same seed frame and timed. Run as `python -m rustracer_amd.convergence`.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import rustracer_amd as rt
from rustracer_amd import _ffi

EPS = 1e-2
THRESHOLD = 0.01


def _log(msg: str) -> None:
    print(f"[convergence] {msg}", file=sys.stderr, flush=True)


def _set(r, **fields) -> None:
    c = r.config
    for k, v in fields.items():
        setattr(c, k, v)
    r.config = c


def build_reference(r, ref_frames: int) -> float:
    """FP64 mean of ref_frames raw 1-spp frames (no a-trous, alpha = 1 so the
    accumulation holds each frame). Returns the seconds it took."""
    _set(r, enable_denoise=False, accumulate_frames=1)
    r.reset_accumulation()
    r.accum64_reset()
    t0 = time.perf_counter()
    for k in range(ref_frames):
        r.record_frame()
        r.accum64_add()
        if (k + 1) % 2048 == 0:
            r.synchronize()
            _log(f"reference {k + 1}/{ref_frames} frames")
    r.synchronize()
    return time.perf_counter() - t0


def run_method(r, name: str, denoise: bool, start_frame: int, max_frames: int, unet=None):
    """First frame count reaching the threshold, its timed back-to-back wall
    time, and the relMSE curve at powers of two. With `unet` the displayed
    image is the learned denoiser's output (run every frame)."""
    _set(r, enable_denoise=denoise, accumulate_frames=1 << 20)
    r.set_denoiser(unet)
    r.reset_accumulation()
    r.set_sample_count(start_frame)
    curve, hit = {}, None
    for k in range(1, max_frames + 1):
        r.record_frame()
        e = r.accum64_relmse_display(EPS)
        if k & (k - 1) == 0:
            curve[k] = e
        if e <= THRESHOLD:
            hit = (k, e)
            break
    result = {"method": name, "curve": curve}
    if hit is None:
        result.update(frames=None, relmse_at_end=e, frames_run=max_frames)
        _log(f"{name}: relMSE {e:.4f} after {max_frames} frames (threshold not reached)")
        r.set_denoiser(None)
        return result
    k, e = hit
    r.reset_accumulation()
    r.set_sample_count(start_frame)
    r.synchronize()
    t0 = time.perf_counter()
    r.render_frames(k)  # as a frame loop issues them: up to eight frames per camera-pass launch
    r.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    check = r.accum64_relmse_display(EPS)
    result.update(frames=k, relmse=e, relmse_rerun=check, ms=round(ms, 3))
    _log(f"{name}: relMSE {e:.4f} after {k} frames, {ms:.1f} ms back to back")
    r.set_denoiser(None)
    return result


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--scene", default="fullframe")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--photons", type=int, default=1_000_000)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--ref-frames", type=int, default=16384)
    ap.add_argument("--max-frames", type=int, default=4096)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--train-scenes", type=int, default=64)
    ap.add_argument("--train-steps", type=int, default=800)
    ap.add_argument("--weights", default=None,
                    help="use these trained parameters (float32, from `python -m rustracer_amd.denoiser --out`) "
                         "instead of training now")
    args = ap.parse_args(argv)
    unet, train_info = None, None
    if args.weights and not args.no_learned:
        import numpy as np
        from .denoiser import UNet
        unet = UNet(0, 1)
        unet.set_params(np.fromfile(args.weights, np.float32))
        train_info = {"network": "5-level U-Net, 32-128 channels, 12 input channels, L1 on log radiance",
                      "weights": args.weights,
                      "trained_by": "python -m rustracer_amd.denoiser --scenes 2000 --steps 2000 --out <weights> "
                                    "(for the shipped rustracer_amd/weights/unet_config4.f32)"}
    elif not args.no_learned:
        from .denoiser import train_denoiser
        unet, losses, (data_s, train_s) = train_denoiser(args.train_scenes, args.train_steps, log=_log)
        train_info = {"network": "5-level U-Net, 32-128 channels, 12 input channels, L1 on log radiance",
                      "scenes": args.train_scenes, "steps": args.train_steps, "batch": "32 x 256x256",
                      "inputs": "1/2/4/8-frame accumulations", "target_samples": 4096,
                      "data_seconds": round(data_s, 2),
                      "train_seconds": round(train_s, 2), "loss_first10": float(losses[:10].mean()),
                      "loss_last10": float(losses[-10:].mean())}
    scene = rt.Scene.builtin(args.scene, 1)
    cam = rt.Scene.builtin_camera(args.scene, 1)
    cfg = rt.RenderConfig(width=args.width, height=args.height, spp=1, max_depth=args.depth,
                          enable_photon_map=args.photons > 0, photon_count=max(args.photons, 1))
    with rt.Renderer(cfg, scene, 0) as r:
        r.update_camera(cam)
        if args.photons > 0:
            r.rebuild_photons()
        ref_s = build_reference(r, args.ref_frames)
        _log(f"reference: {args.ref_frames} frames in {ref_s:.1f} s")
        start = r.sample_count() + 1000
        out = {"scene": args.scene, "width": args.width, "height": args.height, "photons": args.photons,
               "depth": args.depth, "eps": EPS, "threshold": THRESHOLD, "reference_spp": args.ref_frames,
               "reference_seconds": round(ref_s, 2), "methods": []}
        out["methods"].append(run_method(r, "a-trous x2 + running mean", True, start, args.max_frames))
        out["methods"].append(run_method(r, "running mean (no filter)", False, start, args.max_frames))
        if unet is not None:
            out["training"] = train_info
            out["methods"].append(run_method(r, "learned U-Net on the running mean", False, start,
                                             args.max_frames, unet))
            base = out["methods"][0]
            learned = out["methods"][-1]
            if base.get("ms") and learned.get("ms"):
                out["speedup_vs_atrous"] = round(base["ms"] / learned["ms"], 3)
    print(json.dumps(out), flush=True)
    if unet is not None:
        unet.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
