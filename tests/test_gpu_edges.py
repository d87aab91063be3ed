"""GPU edge cases against the oracle: an empty scene, images smaller than one
tile and ragged sizes under every camera-pass form, the depth extremes, and a
BVH deeper than the reference's 63-entry stack (the fallback traversal must
drop the same subtrees the reference drops)."""
import numpy as np
import pytest

import rustracer_amd as rt
from rustracer_amd import _ffi

from . import oracle_ffi as orc
from .helpers import image_params
from .test_oracle_trace import _chain_bvh

pytestmark = pytest.mark.gpu


def _equal(name, a, b):
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name


def test_empty_scene_renders_black_with_default_gbuffer():
    s = rt.Scene()
    s.add_light(_ffi.LIGHT_ENVIRONMENT, vec_a=(0.4, 0.5, 0.6), vec_b=(0.9, 0.8, 0.7), intensity=2.0)
    cam = rt.camera((0.0, 1.0, 3.0), (0.0, 0.0, 0.0))
    print("empty scene", flush=True)
    with rt.Renderer(rt.RenderConfig(width=33, height=17, spp=2, enable_denoise=True, accumulate_frames=4), s,
                     0) as r:
        r.update_camera(cam)
        r.render_frames(2)
        frame, gb, out = r.read_frame(), r.read_gbuffer(), r.read_output_rgba8()
        cfg = r.config
    ref = orc.OracleScene.from_scene(s).gather(rt.camera_to_uniform(cam, 33, 17), image_params(cfg, 2))
    assert (frame[..., :3] == 0).all(), "primary misses see black"
    _equal("gbuffer", gb, ref[2])
    assert (gb[..., 3] == np.float32(1e30)).all()
    assert (out[..., :3] == 0).all() and (out[..., 3] == 255).all()


@pytest.mark.parametrize("flags", [0, _ffi.FLAG_TILE_LAUNCH, _ffi.FLAG_PATH_QUEUES])
@pytest.mark.parametrize("w,h", [(1, 1), (7, 3), (9, 65), (130, 2)])
def test_small_and_ragged_images(w, h, flags):
    scene = rt.Scene.builtin("cornell_box", 2)
    cam = rt.Scene.builtin_camera("cornell_box", 2)
    print(f"image {w}x{h} flags={flags}", flush=True)
    with rt.Renderer(rt.RenderConfig(width=w, height=h, spp=3, enable_denoise=False, flags=flags), scene, 0) as r:
        r.update_camera(cam)
        r.render_frames(2)
        got = (r.read_frame(), r.read_moments(), r.read_gbuffer())
        cfg = r.config
    ref = orc.OracleScene.from_scene(scene).gather(rt.camera_to_uniform(cam, w, h), image_params(cfg, 2))
    for name, a, b in zip(("frame", "moments", "gbuffer"), got, ref[:3]):
        _equal(name, a, b)


@pytest.mark.parametrize("flags", [0, _ffi.FLAG_PATH_QUEUES])
@pytest.mark.parametrize("depth", [1, 2, 64])
def test_depth_extremes(depth, flags):
    """Depth 1 (primary only), 2 and 64 (with path queues: one pass per depth,
    63 of them, most finding their queue empty)."""
    scene = rt.Scene.builtin("photon_plain", 2)
    cam = rt.Scene.builtin_camera("photon_plain", 2)
    w, h = 40, 24
    print(f"max_depth={depth} flags={flags}", flush=True)
    with rt.Renderer(rt.RenderConfig(width=w, height=h, spp=2, max_depth=depth, enable_denoise=False, flags=flags),
                     scene, 0) as r:
        r.update_camera(cam)
        r.render_frames(1)
        got = (r.read_frame(), r.read_moments(), r.read_gbuffer())
        cfg = r.config
    ref = orc.OracleScene.from_scene(scene).gather(rt.camera_to_uniform(cam, w, h), image_params(cfg, 1),
                                                   max_depth=depth)
    for name, a, b in zip(("frame", "moments", "gbuffer"), got, ref[:3]):
        _equal(name, a, b)


@pytest.mark.parametrize("depth,lvl,visible", [(40, 39, True), (70, 62, True), (70, 63, False)])
def test_bvh_deeper_than_the_reference_stack(depth, lvl, visible):
    """A 70-deep chain: the child-box layouts are refused (the reference would
    drop subtrees), the reference-order traversal runs, and the right leaf of
    level 63 is dropped exactly as gather.wgsl:84 drops it."""
    tris, nodes = _chain_bvh(depth)
    prims = np.arange(len(tris), dtype=np.uint32)
    prims[[0, lvl + 1]] = prims[[lvl + 1, 0]]
    mats = rt.make_material(0)
    lights = np.zeros(1, _ffi.LIGHT)
    lights[0]["kind"] = _ffi.LIGHT_ENVIRONMENT
    bounds = (np.array([-200, -200, -200], np.float32), np.array([200, 200, 200], np.float32))
    ray = np.zeros(4, _ffi.RAY)
    ray["dir"] = (0, 0, 1)
    ray["origin"][1] = (0.1, 0.0, 0.0)
    ray["origin"][2] = (0.0, 0.2, 0.0)
    ray["dir"][3] = (0.0, 0.0, -1.0)
    print(f"chain depth={depth} lvl={lvl}", flush=True)
    with rt.Renderer(rt.RenderConfig(width=8, height=8), arrays=(tris, nodes, prims, mats, lights, bounds)) as r:
        info = r.traversal_info()
        got = r.trace_closest(ray)
    if depth > 63:
        assert info["mode"] == 0, info
        # the reference traversal defers at most 63 one-word entries (gather.wgsl:84)
        assert info["stack_words"] == 63, info
    ref = orc.OracleScene(tris, nodes, prims, mats, lights).trace(ray)
    assert np.array_equal(got["t"].view(np.uint32), ref["t"].view(np.uint32))
    hit = got["t"] < 1e38
    assert np.array_equal(got["prim_id"][hit], ref["prim_id"][hit])
    assert (got["t"][0] == 5.0) == visible


@pytest.mark.parametrize("scene_name,photons", [("cornell_box", False), ("photon_plain", True)])
def test_one_sample_frames_bit_exact(scene_name, photons):
    """1 spp (the persistent pass keeps a one-sample accumulator for it): frame,
    moments and G-buffer equal the oracle's, including the firefly clamp's
    one-sample window."""
    scene = rt.Scene.builtin(scene_name, 3)
    cam = rt.Scene.builtin_camera(scene_name, 3)
    w, h = 61, 37
    print(f"1 spp {scene_name}", flush=True)
    cfg_kw = dict(width=w, height=h, spp=1, max_depth=8, enable_denoise=False, enable_photon_map=photons,
                  photon_count=30000)
    with rt.Renderer(rt.RenderConfig(**cfg_kw), scene, 0) as r:
        r.update_camera(cam)
        if photons:
            r.rebuild_photons()
        r.render_frames(3)
        got = (r.read_frame(), r.read_moments(), r.read_gbuffer())
        cfg = r.config
        pm = (r.grid_params(), r.read_grid_meta(), r.read_sorted_photons()) if photons else None
    o = orc.OracleScene.from_scene(scene)
    ref = o.gather(rt.camera_to_uniform(cam, w, h), image_params(cfg, 3), max_depth=8,
                   photon_map=pm)
    for name, a, b in zip(("frame", "moments", "gbuffer"), got, ref[:3]):
        _equal(name, a, b)
