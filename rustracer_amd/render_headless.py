"""render_headless with the argument set of rustracer-app/examples/render_headless.rs:

    python -m rustracer_amd.render_headless <scene> [frames] [out.png] [denoise] [app_like]
        [photons] [spp] [accum] [w] [h] [debug_view] [zoom] [yaw] [pitch] [raw_save]

<scene> is a .gltf / .glb file (with a sibling scene.toml applied when present,
render_headless.rs:68-76) or a built-in scene name with an optional ":seed".
Defaults follow render_headless.rs:18-114: 128 frames, photon_count 50 000,
app_like = spp 4 + accumulation 32 + denoise, accumulation defaulting to the
frame count, Camera::frame_scene then zoom and orbit. The accumulation buffer is
saved as a PNG with the CPU tonemap of examples/common/mod.rs:37-47, or as raw
little-endian f32 RGBA in <out>.bin with raw_save.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import sys

import numpy as np

from . import RenderConfig, Renderer, Scene, _ffi
from ._ffi import check, lib, ptr


# The grammars of Rust's str::parse::<u32>() and str::parse::<f32>(): no
# whitespace, no underscores, no hex; a u32 takes an optional '+' and at most
# u32::MAX; an f32 takes [+-] then inf / infinity / nan (any case) or a decimal
# number with an optional exponent.
_U32 = re.compile(r"\+?[0-9]+")
_F32 = re.compile(r"[+-]?(?:inf|infinity|nan|(?:[0-9]+(?:\.[0-9]*)?|\.[0-9]+)(?:e[+-]?[0-9]+)?)", re.I)


def parse_u32(s: str):
    if _U32.fullmatch(s) and int(s) <= 0xFFFFFFFF:
        return int(s)
    return None


def parse_f32(s: str):
    if _F32.fullmatch(s):
        return float(np.float32(float(s)))
    return None


def _arg(argv, i, parse, default):
    """render_headless.rs parses with .parse().ok(): a value that does not parse keeps the default."""
    if len(argv) > i:
        v = parse(argv[i])
        return default if v is None else v
    return default


def load_scene(arg: str) -> Scene:
    if arg.endswith((".gltf", ".glb")):
        from .gltf import load_gltf

        scene = load_gltf(arg)
        desc = os.path.join(os.path.dirname(os.path.abspath(arg)), "scene.toml")
        if os.path.exists(desc):
            scene.apply_descriptor_file(desc)
            print(f"applied {desc}")
        return scene
    name, _, seed = arg.partition(":")
    return Scene.builtin(name, int(seed) if seed else 1)


def tonemap_cpu(acc: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(acc, np.float32)
    out = np.zeros(a.shape[:-1] + (4,), np.uint8)
    check(lib().rt_tonemap_cpu(ptr(a), a.size // 4, ptr(out)), "rt_tonemap_cpu")
    return out


def main(argv=None) -> int:
    argv = list(sys.argv if argv is None else argv)
    if len(argv) < 2:
        print(__doc__.strip().splitlines()[2].strip(), file=sys.stderr)
        return 2
    frames = _arg(argv, 2, parse_u32, 128)
    out = argv[3] if len(argv) > 3 else "render_headless.png"
    flag = lambda i: len(argv) > i and argv[i] == "1"
    denoise, app_like, photons = flag(4), flag(5), flag(6)
    spp = _arg(argv, 7, parse_u32, 4 if app_like else 1)
    accum = _arg(argv, 8, parse_u32, 32 if app_like else frames)
    w, h = _arg(argv, 9, parse_u32, 800), _arg(argv, 10, parse_u32, 600)
    debug_view = _arg(argv, 11, parse_u32, 0)
    zoom, yaw, pitch = _arg(argv, 12, parse_f32, 0.0), _arg(argv, 13, parse_f32, 0.0), _arg(argv, 14, parse_f32, 0.0)
    raw_save = flag(15)

    scene = load_scene(argv[1])
    c = scene.counts()
    print(f"scene: {c['meshes']} meshes, {c['materials']} mats, {c['triangles']} tris, {c['lights']} lights")
    cfg = RenderConfig(width=w, height=h, spp=spp, accumulate_frames=accum, enable_denoise=denoise or app_like,
                       enable_photon_map=photons, photon_count=50_000, debug_view=debug_view)
    with Renderer(cfg, scene, 0) as r:
        if photons:
            r.rebuild_photons()
        cam = _ffi.Camera()
        check(lib().rt_camera_frame_scene(scene.handle, C.byref(cam)), "rt_camera_frame_scene")
        if zoom != 0.0:
            check(lib().rt_camera_zoom(C.byref(cam), zoom), "rt_camera_zoom")
        if yaw != 0.0 or pitch != 0.0:
            check(lib().rt_camera_orbit(C.byref(cam), yaw, pitch), "rt_camera_orbit")
        r.update_camera(cam)
        r.render_frames(frames)
        r.synchronize()
        acc = r.read_accumulation()
    if raw_save:
        path = out + ".bin"
        acc.astype("<f4").tofile(path)
        print(f"saved raw {path}")
    else:
        from PIL import Image

        Image.fromarray(tonemap_cpu(acc)).save(out)
        print(f"saved {out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
