"""GPU parity: the gfx950 kernels through the C ABI against the CPU oracle.

Integer/index results (hit primitive, photon-grid offsets, sorted photon order,
rgba8) must match exactly; float outputs are compared bit for bit as well
(tolerance 0): the kernels and the oracle evaluate the same WGSL arithmetic
with the same builtin definitions and no FMA contraction.
"""
import ctypes
import os

import numpy as np
import pytest

import rustracer_amd as rt
from rustracer_amd import _ffi

from . import oracle_ffi as orc
from .helpers import camera_rays, face_rays, first_diff, image_params, oracle_frames, random_rays

pytestmark = pytest.mark.gpu

FLAG_VARIANTS = [0, _ffi.FLAG_REFERENCE_TRAVERSAL, _ffi.FLAG_CLOSEST_SHADOW,
                 _ffi.FLAG_REFERENCE_TRAVERSAL | _ffi.FLAG_CLOSEST_SHADOW, _ffi.FLAG_BINARY_RECORDS,
                 _ffi.FLAG_PATH_QUEUES, _ffi.FLAG_TILE_LAUNCH]
# traversal each flag set selects: 3 = 4-wide records, 1 = binary records, 0 = BVH2 nodes
EXPECTED_MODE = {0: 3, _ffi.FLAG_BINARY_RECORDS: 1, _ffi.FLAG_REFERENCE_TRAVERSAL: 0}


def _renderer(scene, **cfg):
    c = rt.RenderConfig(**cfg)
    return rt.Renderer(c, scene, 0)


def _check_equal(name, a, b):
    eq = (a == b) | (np.isnan(a) & np.isnan(b))
    assert eq.all(), f"{name}: {first_diff(a, b)}"


@pytest.mark.parametrize("scene_name", ["cornell_box", "photon_plain", "multiview"])
@pytest.mark.parametrize("flags", [0, _ffi.FLAG_BINARY_RECORDS, _ffi.FLAG_REFERENCE_TRAVERSAL])
def test_trace_closest_bit_exact(scene_name, flags):
    scene = rt.Scene.builtin(scene_name, 3)
    o = orc.OracleScene.from_scene(scene)
    lo, hi = scene.bounds()
    rays = np.concatenate([camera_rays(rt.Scene.builtin_camera(scene_name, 3), 96, 64, 1),
                           random_rays(lo, hi, 20000, 7)])
    print(f"trace {scene_name} flags={flags} rays={len(rays)}", flush=True)
    with _renderer(scene, width=8, height=8, flags=flags) as r:
        assert r.traversal_info()["mode"] == EXPECTED_MODE[flags]
        got = r.trace_closest(rays)
    ref = o.trace(rays)
    brute = o.brute(rays)
    _check_equal("t", got["t"], ref["t"])
    assert np.array_equal(got["prim_id"], ref["prim_id"]), first_diff(got["prim_id"], ref["prim_id"])
    # the reference's BVH answer is the brute-force closest hit except where a
    # hit's rounded t lies before its own box's entry (the strict pruning then
    # depends on the visiting order); it is never nearer than brute force
    assert (ref["t"] >= brute["t"]).all()
    assert (ref["t"] == brute["t"]).mean() > 0.999
    assert (ref["t"] < 1e38).sum() > len(rays) // 10


@pytest.mark.parametrize("scene_name", ["cornell_box", "cornell36", "photon_plain"])
@pytest.mark.parametrize("flags", [0, _ffi.FLAG_BINARY_RECORDS, _ffi.FLAG_REFERENCE_TRAVERSAL])
def test_trace_face_plane_rays_bit_exact(scene_name, flags):
    """Rays lying in box-face planes with zero direction components: the slab
    test sees 0 * inf = NaN, flat boxes (axis-aligned walls) ignore the axis and
    non-flat boxes with a face in the plane fail. The 4-wide walk skips
    intermediate boxes and must still take every decision the reference takes."""
    scene = rt.Scene.builtin(scene_name, 3)
    nodes, _ = scene.build_bvh()
    rays = face_rays(nodes, 30000, 21)
    print(f"face rays {scene_name} flags={flags} rays={len(rays)}", flush=True)
    with _renderer(scene, width=8, height=8, flags=flags) as r:
        assert r.traversal_info()["mode"] == EXPECTED_MODE[flags]
        got = r.trace_closest(rays)
    ref = orc.OracleScene.from_scene(scene).trace(rays)
    _check_equal("t", got["t"], ref["t"])
    assert np.array_equal(got["prim_id"], ref["prim_id"]), first_diff(got["prim_id"], ref["prim_id"])
    assert (ref["t"] < 1e38).sum() > len(rays) // 20


@pytest.mark.parametrize("scene_name", ["cornell_box", "photon_plain"])
def test_trace_nearest_first_mismatch_count(scene_name):
    """RT_FLAG_NEAREST_FIRST is not exact by construction: count its differences
    from the reference order (expected: only rounding-boundary ties)."""
    scene = rt.Scene.builtin(scene_name, 3)
    lo, hi = scene.bounds()
    rays = np.concatenate([camera_rays(rt.Scene.builtin_camera(scene_name, 3), 96, 64, 1),
                           random_rays(lo, hi, 20000, 7)])
    with _renderer(scene, width=8, height=8, flags=_ffi.FLAG_NEAREST_FIRST) as r:
        got = r.trace_closest(rays)
    ref = orc.OracleScene.from_scene(scene).trace(rays)
    bad = (got["t"] != ref["t"]) | (got["prim_id"] != ref["prim_id"])
    print(f"nearest-first {scene_name}: {bad.sum()} of {len(rays)} rays differ", flush=True)
    assert bad.mean() < 1e-3
    dt = np.abs(got["t"][bad].astype(np.float64) - ref["t"][bad]) / ref["t"][bad]
    assert (dt < 1e-5).all()


@pytest.mark.parametrize("scene_name,spp,depth", [("cornell_box", 4, 6), ("cornell36", 16, 5),
                                                  ("smoke", 2, 6), ("photon_plain", 3, 6)])
@pytest.mark.parametrize("flags", FLAG_VARIANTS)
def test_gather_frame_bit_exact(scene_name, spp, depth, flags):
    scene = rt.Scene.builtin(scene_name, 5)
    cam = rt.Scene.builtin_camera(scene_name, 5)
    w, h = 72, 40
    print(f"gather {scene_name} spp={spp} depth={depth} flags={flags}", flush=True)
    with _renderer(scene, width=w, height=h, spp=spp, max_depth=depth, enable_denoise=False, flags=flags) as r:
        r.update_camera(cam)
        r.render_frames(2)
        got = (r.read_frame(), r.read_moments(), r.read_gbuffer())
        cfg = r.config
    o = orc.OracleScene.from_scene(scene)
    ref = o.gather(rt.camera_to_uniform(cam, w, h), image_params(cfg, 2), max_depth=depth)
    for name, a, b in zip(("frame", "moments", "gbuffer"), got, ref[:3]):
        _check_equal(name, a, b)


def test_cornell_256_64spp_bit_exact():
    """The Cornell box workload at its full shape: 36 triangles, 256x256, 64 spp,
    depth 5, photon map and denoise off; frame, moments and G-buffer against
    the oracle."""
    scene = rt.Scene.builtin("cornell36", 1)
    cam = rt.Scene.builtin_camera("cornell36", 1)
    print("cornell36 256x256 64 spp against the oracle", flush=True)
    with _renderer(scene, width=256, height=256, spp=64, max_depth=5, enable_denoise=False) as r:
        r.update_camera(cam)
        r.record_frame()
        got = (r.read_frame(), r.read_moments(), r.read_gbuffer())
        cfg = r.config
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    ref = orc.OracleScene.from_scene(scene).gather(rt.camera_to_uniform(cam, 256, 256), image_params(cfg, 1),
                                                   max_depth=5, threads=threads)
    for name, a, b in zip(("frame", "moments", "gbuffer"), got, ref[:3]):
        _check_equal(name, a, b)


@pytest.mark.parametrize("view", [1, 2, 4, 13, 17, 19, 20, 21])
def test_debug_views_bit_exact(view):
    scene = rt.Scene.builtin("cornell_box")
    cam = rt.Scene.builtin_camera("cornell_box")
    with _renderer(scene, width=40, height=30, spp=2, debug_view=view, enable_denoise=False) as r:
        r.update_camera(cam)
        r.record_frame()
        got = r.read_frame()
        cfg = r.config
    ref = orc.OracleScene.from_scene(scene).gather(rt.camera_to_uniform(cam, 40, 30), image_params(cfg, 1))
    _check_equal("frame", got, ref[0])


def _photon_setup(n_photons, cell=0.0, seed=11):
    scene = rt.Scene.builtin("photon_plain", seed)
    cfg = dict(width=48, height=32, spp=2, enable_denoise=False, enable_photon_map=True, photon_count=n_photons,
               grid_cell_size=cell)
    return scene, cfg


@pytest.mark.parametrize("cell", [0.0, 0.05])
def test_photon_trace_and_grid_bit_exact(cell):
    n = 4096 + 77
    scene, cfg = _photon_setup(n, cell)
    print(f"photon trace n={n} cell={cell}", flush=True)
    with _renderer(scene, **cfg) as r:
        r.rebuild_photons()
        photons = r.read_photons()
        meta = r.read_grid_meta()
        srt = r.read_sorted_photons()
        gp = r.grid_params()
    o = orc.OracleScene.from_scene(scene)
    traced = (n // 256) * 256
    ref_ph = o.photon_trace(gp, traced, 6, n_slots=gp.photon_slots)
    _check_equal("photons", photons.view(np.float32), ref_ph.view(np.float32))
    n_proc = min(max(n * 6 // 256, 1) * 256, gp.photon_slots)
    ref_meta, ref_sorted = orc.photon_grid(ref_ph, n_proc, gp, len(meta))
    assert np.array_equal(meta, ref_meta), first_diff(meta, ref_meta)
    assert len(srt) == len(ref_sorted) > 0
    _check_equal("sorted photons", srt.view(np.float32), ref_sorted.view(np.float32))


def test_traced_keys_equal_keys_from_records():
    """The photon trace writes each stored photon's cell key as it stores the
    record; the grid built from those keys must equal the grid built again from
    the read-back photon array (keys computed from the records). 300k photons,
    1.8M slots, on the 0.05 grid; then an array with no valid photon and an
    empty array (the build's zero paths)."""
    scene, cfg = _photon_setup(300_000, 0.05)
    print("traced keys vs record keys, 300k photons", flush=True)
    with _renderer(scene, **cfg) as r:
        r.rebuild_photons()
        meta_t, srt_t = r.read_grid_meta(), r.read_sorted_photons()
        ph = r.read_photons()
        r.build_photon_grid_from(ph)
        meta_u, srt_u = r.read_grid_meta(), r.read_sorted_photons()
        assert len(srt_t) > 10_000
        assert np.array_equal(meta_t, meta_u), first_diff(meta_t, meta_u)
        _check_equal("sorted photons", srt_t.view(np.float32), srt_u.view(np.float32))
        dark = ph.copy()
        dark["power"][:, 0] = 0.0
        r.build_photon_grid_from(dark)
        meta_d = r.read_grid_meta()
        assert len(r.read_sorted_photons()) == 0
        assert not meta_d.any()
        r.build_photon_grid_from(ph[:0])
        assert len(r.read_sorted_photons()) == 0
        assert not r.read_grid_meta().any()


def test_grid_build_large_matches_stable_sort():
    """1.5M synthetic photon slots (a quarter empty) on a 0.05 grid: offsets and
    order equal numpy's stable argsort of the cell keys."""
    scene, cfg = _photon_setup(250_000, 0.05)
    rng = np.random.default_rng(5)
    with _renderer(scene, **cfg) as r:
        gp = r.grid_params()
        n = gp.photon_slots
        ph = np.zeros(n, _ffi.PHOTON)
        lo = np.array([gp.room_min_x, gp.room_min_y, gp.room_min_z], np.float32)
        span = np.array([gp.grid_x, gp.grid_y, gp.grid_z], np.float32) * np.float32(gp.cell_size)
        ph["position"] = (lo + rng.random((n, 3), np.float32) * span * np.float32(1.02)).astype(np.float32)
        ph["power"] = rng.random((n, 3), np.float32) + np.float32(0.01)
        ph["power"][rng.random(n) < 0.25, 0] = 0.0
        ph["incident"] = rng.normal(size=(n, 3)).astype(np.float32)
        r.build_photon_grid_from(ph)
        meta = r.read_grid_meta()
        srt = r.read_sorted_photons()
    # keys exactly as photon_count.wgsl:18-23 computes them, in float32
    rel = (ph["position"] - lo) / np.float32(gp.cell_size)
    g = np.where(rel > 0, np.minimum(rel, np.float32(4294967040.0)), 0).astype(np.uint64)
    key = g[:, 0] + g[:, 1] * gp.grid_x + g[:, 2] * gp.grid_x * gp.grid_y
    key = key.astype(np.uint32)  # u32 wrap
    valid = (ph["power"][:, 0] > 0) & (key < gp.num_cells)
    order = np.argsort(np.where(valid, key, gp.num_cells), kind="stable")[: valid.sum()]
    counts = np.bincount(key[valid], minlength=len(meta))
    assert np.array_equal(meta[:, 1], counts)
    assert np.array_equal(meta[:, 0], np.concatenate([[0], np.cumsum(counts)[:-1]]))
    assert np.array_equal(srt.view(np.float32).reshape(-1, 16), ph[order].view(np.float32).reshape(-1, 16))


def test_gather_with_photon_map_bit_exact():
    scene, cfg = _photon_setup(20000, 0.05)
    cam = rt.Scene.builtin_camera("photon_plain", 11)
    with _renderer(scene, **cfg) as r:
        r.update_camera(cam)
        r.rebuild_photons()
        r.record_frame()
        got = (r.read_frame(), r.read_moments(), r.read_gbuffer())
        gp, meta, srt, c = r.grid_params(), r.read_grid_meta(), r.read_sorted_photons(), r.config
    o = orc.OracleScene.from_scene(scene)
    ref = o.gather(rt.camera_to_uniform(cam, c.width, c.height), image_params(c, 1), photon_map=(gp, meta, srt))
    for name, a, b in zip(("frame", "moments", "gbuffer"), got, ref[:3]):
        _check_equal(name, a, b)


@pytest.mark.parametrize("denoise,accum,passes", [(True, 0, 2), (True, 4, 2), (False, 3, 2), (True, 5, 3)])
def test_full_pipeline_bit_exact(denoise, accum, passes):
    """gather -> a-trous -> composite -> tonemap over 4 frames (record_frame x 4)."""
    scene = rt.Scene.builtin("cornell_box")
    cam = rt.Scene.builtin_camera("cornell_box")
    w, h = 56, 44
    with _renderer(scene, width=w, height=h, spp=2, enable_denoise=denoise, accumulate_frames=accum,
                   denoise_passes=passes) as r:
        r.update_camera(cam)
        r.render_frames(4)
        acc, rgba = r.read_accumulation(), r.read_output_rgba8()
        cfg = r.config
    o = orc.OracleScene.from_scene(scene)
    racc, rrgba, *_ = oracle_frames(o, cfg, rt.camera_to_uniform(cam, w, h), 4)
    _check_equal("accumulation", acc, racc)
    assert np.array_equal(rgba, rrgba), first_diff(rgba, rrgba)


@pytest.mark.parametrize("flags", [_ffi.FLAG_GRAPH, _ffi.FLAG_GRAPH | _ffi.FLAG_TILE_LAUNCH])
def test_frame_graph_pipeline_matches_oracle(flags):
    """Frames replayed from captured hipGraphs against the oracle (not only
    against direct launches): gather -> a-trous -> composite -> tonemap over
    six frames issued one at a time, the graphs captured at the first frame of
    each output set and replayed after."""
    scene = rt.Scene.builtin("cornell_box", 4)
    cam = rt.Scene.builtin_camera("cornell_box", 4)
    w, h = 48, 40
    print(f"frame graphs against the oracle flags={flags}", flush=True)
    with _renderer(scene, width=w, height=h, spp=2, enable_denoise=True, accumulate_frames=6, flags=flags) as r:
        r.update_camera(cam)
        for _ in range(6):
            r.record_frame()
        acc, rgba = r.read_accumulation(), r.read_output_rgba8()
        cfg = r.config
        captures = ctypes.c_uint32()
        assert _ffi.lib().rt_frame_graph_captures(r._h, ctypes.byref(captures)) == 0
    assert captures.value > 0
    o = orc.OracleScene.from_scene(scene)
    racc, rrgba, *_ = oracle_frames(o, cfg, rt.camera_to_uniform(cam, w, h), 6)
    _check_equal("accumulation", acc, racc)
    assert np.array_equal(rgba, rrgba), first_diff(rgba, rrgba)


def test_fullframe_1080p_properties():
    """The 1M-triangle 1080p frame at full size: every pixel finite, moments
    consistent (E[l^2] >= E[l]^2 up to the clamps), G-buffer normals unit length."""
    scene = rt.Scene.builtin("fullframe", 1)
    cam = rt.Scene.builtin_camera("fullframe", 1)
    with _renderer(scene, width=1920, height=1080, spp=1, max_depth=8, enable_denoise=True) as r:
        r.update_camera(cam)
        r.render_frames(2)
        f, m, g = r.read_frame(), r.read_moments(), r.read_gbuffer()
        out = r.read_output_rgba8()
    assert np.isfinite(f).all() and np.isfinite(m).all()
    # the gather clamps to 10; the a-trous weighted mean may round one ulp above
    assert (f[..., :3] >= 0).all() and (f[..., :3] <= np.float32(10.00001)).all()
    assert (m[..., 0] <= 10).all() and (m[..., 1] <= 100).all()
    hit = g[..., 3] < 1e29
    assert hit.mean() > 0.5
    n = np.linalg.norm(g[..., :3][hit], axis=-1)
    assert np.abs(n - 1).max() < 1e-5
    assert (out[..., 3] == 255).all()


def test_fullframe_1080p_photon_map_bit_exact():
    """The 1080p workload at its full shape with the photon map on (1M photons,
    depth 8, 1M triangles): the photon records, the grid and the sorted
    photons, then the gather's frame colour, moments and G-buffer, bit for bit
    against the oracle on the same inputs."""
    scene = rt.Scene.builtin("fullframe", 1)
    cam = rt.Scene.builtin_camera("fullframe", 1)
    n = 1_000_000
    print("fullframe 1080p with 1M photons against the oracle", flush=True)
    with _renderer(scene, width=1920, height=1080, spp=1, max_depth=8, enable_denoise=False,
                   enable_photon_map=True, photon_count=n) as r:
        r.update_camera(cam)
        r.rebuild_photons()
        photons, meta, srt, gp = r.read_photons(), r.read_grid_meta(), r.read_sorted_photons(), r.grid_params()
        r.record_frame()
        got = (r.read_frame(), r.read_moments(), r.read_gbuffer())
        cfg = r.config
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    o = orc.OracleScene.from_scene(scene)
    ref_ph = o.photon_trace(gp, (n // 256) * 256, 8, n_slots=gp.photon_slots, threads=threads)
    _check_equal("photons", photons.view(np.float32), ref_ph.view(np.float32))
    n_proc = min(max(n * 8 // 256, 1) * 256, gp.photon_slots)
    ref_meta, ref_sorted = orc.photon_grid(ref_ph, n_proc, gp, len(meta))
    assert np.array_equal(meta, ref_meta), first_diff(meta, ref_meta)
    _check_equal("sorted photons", srt.view(np.float32), ref_sorted.view(np.float32))
    assert len(srt) > 100_000
    ref = o.gather(rt.camera_to_uniform(cam, 1920, 1080), image_params(cfg, 1), max_depth=8,
                   photon_map=(gp, meta, srt), threads=threads)
    for name, a, b in zip(("frame", "moments", "gbuffer"), got, ref[:3]):
        _check_equal(name, a, b)


def test_photon_plain_720p_pipeline_bit_exact():
    """The photon-mapped plain scene at its full shape (1280x720, 1M photons,
    grid cell = radius = 0.05, 4 spp, depth 6, a-trous x2, 32-frame running
    mean): the grid and sorted photons, then four frames of the whole pipeline
    (accumulation and rgba8), bit for bit against the oracle."""
    scene = rt.Scene.builtin("photon_plain", 1)
    cam = rt.Scene.builtin_camera("photon_plain", 1)
    w, h, n = 1280, 720, 1_000_000
    print("photon_plain 720p with 1M photons against the oracle", flush=True)
    with _renderer(scene, width=w, height=h, spp=4, max_depth=6, enable_photon_map=True, photon_count=n,
                   grid_cell_size=0.05, enable_denoise=True, accumulate_frames=32) as r:
        r.update_camera(cam)
        r.rebuild_photons()
        photons, meta, srt, gp = r.read_photons(), r.read_grid_meta(), r.read_sorted_photons(), r.grid_params()
        r.render_frames(4)
        acc, rgba = r.read_accumulation(), r.read_output_rgba8()
        cfg = r.config
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    o = orc.OracleScene.from_scene(scene)
    ref_ph = o.photon_trace(gp, (n // 256) * 256, 6, n_slots=gp.photon_slots, threads=threads)
    _check_equal("photons", photons.view(np.float32), ref_ph.view(np.float32))
    ref_meta, ref_sorted = orc.photon_grid(ref_ph, min(max(n * 6 // 256, 1) * 256, gp.photon_slots), gp, len(meta))
    assert np.array_equal(meta, ref_meta), first_diff(meta, ref_meta)
    _check_equal("sorted photons", srt.view(np.float32), ref_sorted.view(np.float32))
    racc, rrgba, *_ = oracle_frames(o, cfg, rt.camera_to_uniform(cam, w, h), 4, photon_map=(gp, meta, srt),
                                    threads=threads)
    _check_equal("accumulation", acc, racc)
    assert np.array_equal(rgba, rrgba), first_diff(rgba, rrgba)


def _view_params(cfg, w, h, spp, frame):
    ip = image_params(cfg, frame)
    ip.width, ip.height, ip.spp = w, h, spp
    return ip


@pytest.mark.parametrize("tuning", [None, (16, 16, 4, 4, 0), (16, 4, 12, 2, 0), (32, 8, 8, 1, 1), (8, 8, 3, 1, 1)])
def test_batch_views_bit_exact(tuning):
    """Many views of uneven resolution and spp in one persistent launch: every
    view equals the oracle's gather for that camera, resolution, spp and seed."""
    scene = rt.Scene.builtin("multiview", 2)
    views = [(rt.Scene.builtin_camera("multiview", 2, k), w, h, spp, 3 + k)
             for k, (w, h, spp) in enumerate([(40, 24, 1), (33, 17, 2), (64, 8, 4), (9, 70, 8), (24, 24, 16)])]
    with _renderer(scene, width=8, height=8, max_depth=8, enable_denoise=False) as r:
        r.set_views(views)
        if tuning is None:
            r.render_batch()
        else:
            r.render_batch(*tuning)
        got = [r.read_batch_view(k) for k in range(len(views))]
        cfg = r.config
    o = orc.OracleScene.from_scene(scene)
    for k, (cam, w, h, spp, frame) in enumerate(views):
        print(f"view {k}: {w}x{h} spp={spp} frame={frame}", flush=True)
        ref = o.gather(rt.camera_to_uniform(cam, w, h), _view_params(cfg, w, h, spp, frame), max_depth=8)
        for name, a, b in zip(("frame", "moments", "gbuffer"), got[k], ref[:3]):
            _check_equal(f"view {k} {name}", a, b)


def test_batch_multiview_workload_bit_exact():
    """The many-view workload at its full shape: the 64 seeded views of the
    109k-triangle scene (128x128 .. 1920x1080, spp 1-16, depth 8) in one
    persistent launch with the autotuner's pick, every view against the
    oracle's gather for its camera, resolution, spp and seed."""
    from rustracer_amd.workloads import multiview_views

    scene = rt.Scene.builtin("multiview", 1)
    views = multiview_views(64, seed=1)
    print("64-view batch workload against the oracle", flush=True)
    with _renderer(scene, width=8, height=8, max_depth=8, enable_denoise=False) as r:
        r.set_views(views)
        r.batch_autotune(repeats=1)
        r.render_batch()  # with the tuning the autotuner kept
        got = [r.read_batch_view(k) for k in range(len(views))]
        cfg = r.config
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    o = orc.OracleScene.from_scene(scene)
    for k, (cam, w, h, spp, frame) in enumerate(views):
        ref = o.gather(rt.camera_to_uniform(cam, w, h), _view_params(cfg, w, h, spp, frame), max_depth=8,
                       threads=threads)
        for name, a, b in zip(("frame", "moments", "gbuffer"), got[k], ref[:3]):
            _check_equal(f"view {k} ({w}x{h}, {spp} spp) {name}", a, b)


def test_batch_autotune_and_totals():
    from rustracer_amd.workloads import multiview_views

    scene = rt.Scene.builtin("multiview", 1)
    views = multiview_views(12, seed=1, scale=0.25)
    with _renderer(scene, width=8, height=8, max_depth=8, enable_denoise=False) as r:
        r.set_views(views)
        px, smp = r.batch_totals()
        assert px == sum(w * h for _, w, h, _, _ in views)
        assert smp == sum(w * h * spp for _, w, h, spp, _ in views)
        r.render_batch()
        first = [r.read_batch_view(k)[0] for k in range(len(views))]
        best, ms = r.batch_autotune(repeats=1)
        print(f"autotune best {best} {ms:.3f} ms", flush=True)
        assert best["tile_w"] * best["tile_h"] % 64 == 0 and ms > 0
        r.render_batch()  # with the tuned parameters: same pixels
        for k in range(len(views)):
            assert np.array_equal(r.read_batch_view(k)[0], first[k])


@pytest.mark.parametrize("other", [_ffi.FLAG_PATH_QUEUES, _ffi.FLAG_TILE_LAUNCH])
@pytest.mark.parametrize("scene_name,spp,photons", [("cornell_box", 3, False), ("photon_plain", 2, True)])
def test_camera_pass_forms_same_work(scene_name, spp, photons, other):
    """The path-queue passes, the one-workgroup-per-tile kernel and the default
    persistent kernel run the same per-pixel work: identical outputs and
    identical traversal/photon counters."""
    scene = rt.Scene.builtin(scene_name, 4)
    cam = rt.Scene.builtin_camera(scene_name, 4)
    out = {}
    print(f"camera pass forms {scene_name} spp={spp} flags={other}", flush=True)
    for flags in (0, other):
        # ragged size: edge tiles hold slots outside the image
        cfg = dict(width=97, height=61, spp=spp, max_depth=6, enable_denoise=True, accumulate_frames=4,
                   enable_photon_map=photons, photon_count=20000, flags=flags)
        with _renderer(scene, **cfg) as r:
            r.update_camera(cam)
            if photons:
                r.rebuild_photons()
            r.render_frames(2)
            cnt = r.count_gather_traffic()
            out[flags] = (r.read_frame(), r.read_moments(), r.read_gbuffer(), r.read_accumulation(),
                          r.read_output_rgba8(), cnt)
    a, b = out[0], out[other]
    for name, x, y in zip(("frame", "moments", "gbuffer", "accumulation", "rgba8"), a[:5], b[:5]):
        _check_equal(name, x, y)
    # per-lane work is the same; wave steps, active lanes per wave step and wave
    # cycles depend on which rays share a wave, which is what the forms change,
    # and only the persistent kernel has a camera loop to count
    def per_lane(c):
        out = {k: v for k, v in c.items() if k not in ("wave_steps", "active_lanes", "lane_steps", "wave_cycles")}
        out["lane_steps"] = {k: v for k, v in c["lane_steps"].items() if k != "camera_loop"}
        return out
    assert per_lane(a[5]) == per_lane(b[5]), (a[5], b[5])
    assert a[5]["lane_steps"]["camera_loop"] == a[5]["lane_steps"]["shading"]  # one bounce per live lane step
    assert a[5]["ray_queries"] > 97 * 61 * spp


@pytest.mark.parametrize("flags", [0, _ffi.FLAG_TILE_LAUNCH])
def test_back_to_back_frames_match_one_at_a_time(flags):
    """Frames issued back to back (consecutive gathers overlap on two streams,
    each frame with its own output set) give the same accumulation, output and
    last-frame buffers as frames run one at a time with a sync in between."""
    scene = rt.Scene.builtin("multiview", 2)
    cam = rt.Scene.builtin_camera("multiview", 2, 5)
    res = []
    print(f"back to back flags={flags}", flush=True)
    for sync_each in (False, True):
        cfg = dict(width=640, height=360, spp=1, max_depth=8, enable_denoise=True, accumulate_frames=16,
                   enable_photon_map=True, photon_count=100000, flags=flags)
        with _renderer(scene, **cfg) as r:
            r.update_camera(cam)
            r.rebuild_photons()
            for _ in range(5):
                r.record_frame()
                if sync_each:
                    r.synchronize()
            res.append((r.read_accumulation(), r.read_output_rgba8(), r.read_frame(), r.read_moments(),
                        r.read_gbuffer(), r.sample_count()))
    for name, a, b in zip(("accumulation", "rgba8", "frame", "moments", "gbuffer"), res[0][:5], res[1][:5]):
        _check_equal(name, a, b)
    assert res[0][5] == res[1][5] == 5


@pytest.mark.parametrize("spp,lead", [(1, 0), (1, 1), (2, 0)])
def test_chunked_frames_match_frame_by_frame(spp, lead):
    """rt_render_frames renders up to eight frames per camera-pass launch (the
    launch's pixel counters run over the frames one after another), as many
    as fit before the output-set rotation wraps. After `lead` single frames,
    19 frames rendered that way (lead 0: launches of 8, 8 and 3 frames; lead 1:
    8, 7 and 4) give the same accumulation, output and last-frame buffers as
    the same frames run one at a time with a sync in between."""
    scene = rt.Scene.builtin("multiview", 2)
    cam = rt.Scene.builtin_camera("multiview", 2, 5)
    res = []
    print(f"chunked frames spp={spp} lead={lead}", flush=True)
    for chunked in (True, False):
        cfg = dict(width=321, height=187, spp=spp, max_depth=8, enable_denoise=True, accumulate_frames=16,
                   enable_photon_map=True, photon_count=100000)
        with _renderer(scene, **cfg) as r:
            r.update_camera(cam)
            r.rebuild_photons()
            for _ in range(lead):
                r.record_frame()
            if chunked:
                r.render_frames(19)
            else:
                for _ in range(19):
                    r.record_frame()
                    r.synchronize()
            res.append((r.read_accumulation(), r.read_output_rgba8(), r.read_frame(), r.read_moments(),
                        r.read_gbuffer(), r.sample_count()))
    for name, a, b in zip(("accumulation", "rgba8", "frame", "moments", "gbuffer"), res[0][:5], res[1][:5]):
        _check_equal(name, a, b)
    assert res[0][5] == res[1][5] == 19 + lead


def test_accum64_relmse_matches_numpy():
    """The GPU relMSE reduction equals a float64 numpy evaluation of the same
    formula on the read-back buffers."""
    scene = rt.Scene.builtin("cornell_box", 2)
    cam = rt.Scene.builtin_camera("cornell_box", 2)
    cfg = dict(width=123, height=77, spp=1, enable_denoise=False, accumulate_frames=1)
    with _renderer(scene, **cfg) as r:
        r.update_camera(cam)
        r.accum64_reset()
        for _ in range(6):
            r.record_frame()
            r.accum64_add()
        ref, n = r.accum64_mean()
        c = r.config
        c.accumulate_frames = 1 << 20
        c.enable_denoise = True
        r.config = c
        r.reset_accumulation()
        for _ in range(3):
            r.record_frame()
        got = r.accum64_relmse(1e-2)
        x = r.read_accumulation().astype(np.float64)
    assert n == 6
    eps = float(np.float32(1e-2))  # the entry point takes eps as a float
    want = np.mean((x[..., :3] - ref[..., :3]) ** 2 / (ref[..., :3] ** 2 + eps))
    assert abs(got - want) <= 1e-12 * max(abs(want), 1e-30), (got, want)
    assert got > 0


@pytest.mark.parametrize("flags", [0, _ffi.FLAG_CLOSEST_SHADOW, _ffi.FLAG_REFERENCE_TRAVERSAL,
                                   _ffi.FLAG_BINARY_RECORDS, _ffi.FLAG_PATH_QUEUES, _ffi.FLAG_TILE_LAUNCH])
def test_textured_gather_bit_exact(flags):
    """Every texture role (albedo, metallic-roughness, emissive incl. area-light
    samples, normal map, alpha cutout on camera and shadow rays) against the oracle."""
    scene = rt.Scene.builtin("textured")
    cam = rt.Scene.builtin_camera("textured")
    w, h = 96, 72
    print(f"textured gather flags={flags}", flush=True)
    with _renderer(scene, width=w, height=h, spp=4, max_depth=6, enable_denoise=False, flags=flags) as r:
        r.update_camera(cam)
        r.render_frames(2)
        got = (r.read_frame(), r.read_moments(), r.read_gbuffer())
        cfg = r.config
    ref = orc.OracleScene.from_scene(scene).gather(rt.camera_to_uniform(cam, w, h), image_params(cfg, 2))
    for name, a, b in zip(("frame", "moments", "gbuffer"), got, ref[:3]):
        _check_equal(name, a, b)


@pytest.mark.parametrize("view", [3, 5, 8, 9, 11, 12, 18])
def test_textured_debug_views_bit_exact(view):
    """spp 16: at spp <= 8 the firefly clamp rescales the colour by
    clamped/total luminance, and a debug view adds no luminance, so the
    reference shows it black (gather.wgsl:548-576)."""
    scene = rt.Scene.builtin("textured")
    cam = rt.Scene.builtin_camera("textured")
    with _renderer(scene, width=40, height=30, spp=16, debug_view=view, enable_denoise=False) as r:
        r.update_camera(cam)
        r.record_frame()
        got = r.read_frame()
        cfg = r.config
    ref = orc.OracleScene.from_scene(scene).gather(rt.camera_to_uniform(cam, 40, 30), image_params(cfg, 1))
    _check_equal("frame", got, ref[0])
    if view not in (5, 12):  # the emissive views: the ceiling light is outside this camera's frame
        assert np.abs(got[..., :3]).sum() > 0


def test_textured_photons_and_raw_upload_bit_exact():
    """Photon trace with textures (albedo alone, normal map, cutout) and grid,
    and a renderer built from raw arrays plus uploaded texture arrays."""
    scene = rt.Scene.builtin("textured")
    cfg = dict(width=48, height=32, spp=2, enable_denoise=False, enable_photon_map=True, photon_count=4096 + 33)
    with _renderer(scene, **cfg) as r:
        r.rebuild_photons()
        photons = r.read_photons()
        meta = r.read_grid_meta()
        gp = r.grid_params()
        r.update_camera(rt.Scene.builtin_camera("textured"))
        r.record_frame()
        frame = r.read_frame()
    o = orc.OracleScene.from_scene(scene)
    n = cfg["photon_count"]
    ref_ph = o.photon_trace(gp, (n // 256) * 256, 6, n_slots=gp.photon_slots)
    _check_equal("photons", photons.view(np.float32), ref_ph.view(np.float32))
    n_proc = min(max(n * 6 // 256, 1) * 256, gp.photon_slots)
    ref_meta, _ = orc.photon_grid(ref_ph, n_proc, gp, len(meta))
    assert np.array_equal(meta, ref_meta), first_diff(meta, ref_meta)
    nodes, prims = scene.build_bvh()
    lo, hi = scene.bounds()
    arrays = (scene.flatten_triangles(), nodes, prims, scene.materials(), scene.lights_gpu(), (lo, hi),
              [scene.texture_array(k) for k in range(4)])
    with rt.Renderer(rt.RenderConfig(**cfg), None, 0, arrays=arrays) as r2:
        r2.rebuild_photons()
        r2.update_camera(rt.Scene.builtin_camera("textured"))
        r2.record_frame()
        _check_equal("raw-upload frame", r2.read_frame(), frame)


@pytest.mark.parametrize("extra", [0, _ffi.FLAG_PATH_QUEUES, _ffi.FLAG_TILE_LAUNCH])
def test_frame_graphs_replay_bit_exact(extra):
    """Frames replayed from captured hipGraphs give the same bits as frames
    launched directly: every replay is checked against the direct run, with
    reads (synchronizing) between some replays and back-to-back replays
    between others, across a camera move with a soft reset and a config
    change. Graphs are captured once per output set, again only for the
    passes whose inputs changed."""
    scene = rt.Scene.builtin("photon_plain", 3)
    cam0 = rt.Scene.builtin_camera("photon_plain", 3)
    cam1 = rt.Scene.builtin_camera("photon_plain", 3)
    _ffi.lib().rt_camera_orbit(ctypes.byref(cam1), 0.3, -0.1)
    print(f"frame graphs flags={extra}", flush=True)
    runs, captures = {}, None
    for flags in (extra, extra | _ffi.FLAG_GRAPH):
        cfg = dict(width=80, height=48, spp=2, max_depth=6, enable_denoise=True, accumulate_frames=8,
                   enable_photon_map=True, photon_count=20000, flags=flags)
        seq = []
        with _renderer(scene, **cfg) as r:
            r.update_camera(cam0)
            r.rebuild_photons()
            for k in range(20):
                if k == 3:
                    r.update_camera(cam1)
                    r.reset_accumulation_soft()
                if k == 17:
                    c = r.config
                    c.exposure = 1.7
                    r.config = c
                r.record_frame()
                if k % 3 != 1:
                    seq.append((r.read_accumulation(), r.read_output_rgba8(), r.read_frame(), r.read_moments(),
                                r.read_gbuffer()))
            if flags & _ffi.FLAG_GRAPH:
                captures = r.frame_graph_captures()
        runs[flags] = seq
    a, b = runs[extra], runs[extra | _ffi.FLAG_GRAPH]
    assert len(a) == len(b) == 13
    for k, (x, y) in enumerate(zip(a, b)):
        for name, p, q in zip(("accumulation", "rgba8", "frame", "moments", "gbuffer"), x, y):
            _check_equal(f"read {k} {name}", p, q)
    # 16 output sets x (gather + post) at first use; frames 17-19 after the
    # exposure change run on sets 1-3, whose post graphs are captured again
    assert captures == 35, captures


@pytest.mark.parametrize("glow", [1.0, 2.5])
def test_light_marker_tonemap_bit_exact(glow):
    """The light gizmo drawn by the tonemap pass (tonemap.wgsl:32-47) at the
    position project_marker gives (lib.rs:865-885): the output equals the
    oracle's tonemap with the oracle's projection, the gizmo changes pixels,
    and hiding it restores the plain tonemap."""
    scene = rt.Scene.builtin("cornell_box")
    cam = rt.Scene.builtin_camera("cornell_box")
    w, h = 64, 48
    pos = tuple(0.7 * np.float32(a) + 0.3 * np.float32(b) for a, b in zip(cam.position, cam.look_at))
    print(f"light marker glow={glow}", flush=True)
    with _renderer(scene, width=w, height=h, spp=2, enable_denoise=False) as r:
        r.update_camera(cam)
        r.set_light_marker(pos, 0.08)
        r.set_light_glow(glow)
        r.render_frames(2)
        acc, gb, rgba = r.read_accumulation(), r.read_gbuffer(), r.read_output_rgba8()
        cfg = r.config
        r.set_light_marker(None)
        r.record_frame()
        acc2, gb2, rgba2 = r.read_accumulation(), r.read_gbuffer(), r.read_output_rgba8()
    marker = orc.project_marker(cam, w, h, pos, 0.08)
    assert marker[3] > 0, "the test point must be in view"
    tm = _ffi.ToneMapUniform(cfg.exposure, cfg.gamma, w, glow, tuple(float(x) for x in marker))
    ref = orc.tonemap(acc, gb, tm)
    assert np.array_equal(rgba, ref), first_diff(rgba, ref)
    plain = orc.tonemap(acc, gb, _ffi.ToneMapUniform(cfg.exposure, cfg.gamma, w, glow, (0.0, 0.0, 0.0, 0.0)))
    assert (rgba != plain).any()
    ref2 = orc.tonemap(acc2, gb2, _ffi.ToneMapUniform(cfg.exposure, cfg.gamma, w, glow, (0.0, 0.0, 0.0, 0.0)))
    assert np.array_equal(rgba2, ref2)


def test_gltf_scene_gather_bit_exact(tmp_path):
    """A scene loaded from a glTF file (textured floor, emissive-textured panel
    as area lights, a BLEND material rendered as glass, node transforms) renders
    bit-exactly as the oracle renders the same loaded scene."""
    from rustracer_amd import gltf as G

    from .test_gltf import Builder

    rng = np.random.default_rng(11)
    b = Builder()
    floor_tex = b.png_texture(rng.integers(0, 256, (8, 8, 4), dtype=np.uint8))
    glow = np.zeros((8, 8, 4), np.uint8)
    glow[..., 3] = 255
    glow[:4, :, :3] = (255, 200, 120)
    glow_tex = b.png_texture(glow)
    floor = b.material(pbrMetallicRoughness={"baseColorTexture": {"index": floor_tex}, "metallicFactor": 0.0})
    panel = b.material(pbrMetallicRoughness={"baseColorFactor": [0.2, 0.2, 0.2, 1.0]},
                       emissiveTexture={"index": glow_tex}, emissiveFactor=[4.0, 4.0, 4.0])
    glass = b.material(alphaMode="BLEND", pbrMetallicRoughness={"baseColorFactor": [0.9, 0.95, 1.0, 0.3]})
    quad = [[-1, 0, -1], [1, 0, -1], [-1, 0, 1], [1, 0, 1]]
    uvs = [[0, 0], [2, 0], [0, 2], [2, 2]]
    b.node(mesh=b.mesh(b.primitive(quad, [0, 2, 1, 1, 2, 3], uvs=uvs, material=floor)), scale=[3, 1, 3])
    b.node(mesh=b.mesh(b.primitive(quad, [0, 1, 2, 1, 3, 2], uvs=[[0, 0], [1, 0], [0, 1], [1, 1]],
                                   material=panel)), translation=[0, 2.0, 0], scale=[0.5, 1, 0.5])
    tent = [[0, 0, 0], [0.5, 0, 0], [0.25, 0.6, 0.1], [0.25, 0, 0.5]]
    b.node(mesh=b.mesh(b.primitive(tent, [0, 2, 1, 0, 3, 2, 1, 2, 3, 0, 1, 3], material=glass)),
           translation=[-0.2, 0.0, 0.2])
    s = G.load_gltf(b.save(tmp_path / "scene.gltf"))
    assert sum(l["kind"] == _ffi.LIGHT_AREA for l in s.light_descs()) >= 1
    cam = rt.camera((0.0, 1.2, 2.6), (0.0, 0.3, 0.0))
    w, h = 64, 48
    print("gltf scene gather", flush=True)
    with _renderer(s, width=w, height=h, spp=4, max_depth=6, enable_denoise=False) as r:
        r.update_camera(cam)
        r.render_frames(2)
        got = (r.read_frame(), r.read_moments(), r.read_gbuffer())
        cfg = r.config
    o = orc.OracleScene.from_scene(s)
    ref = o.gather(rt.camera_to_uniform(cam, w, h), image_params(cfg, 2), max_depth=6)
    for name, a, bb in zip(("frame", "moments", "gbuffer"), got, ref[:3]):
        _check_equal(name, a, bb)
    assert got[0][..., :3].max() > 0


def test_python_render_headless_on_gltf(tmp_path):
    """`python -m rustracer_amd.render_headless file.gltf ...` writes the PNG of
    the accumulation that the same configuration renders in this process."""
    import subprocess
    import sys

    from PIL import Image

    from rustracer_amd import gltf as G
    from rustracer_amd.render_headless import tonemap_cpu

    from .test_gltf import Builder

    b = Builder()
    quad = [[-1, 0, -1], [1, 0, -1], [-1, 0, 1], [1, 0, 1]]
    mat = b.material(pbrMetallicRoughness={"baseColorFactor": [0.7, 0.5, 0.3, 1], "metallicFactor": 0.0})
    b.node(mesh=b.mesh(b.primitive(quad, [0, 2, 1, 1, 2, 3], material=mat)))
    b.node(mesh=b.mesh(b.primitive([[0, 0, 0], [0.5, 0, 0], [0, 0.7, 0]], [0, 1, 2], material=mat)))
    path = b.save(tmp_path / "cli.gltf")
    out = str(tmp_path / "cli.png")
    print("python render_headless", flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    res = subprocess.run([sys.executable, "-m", "rustracer_amd.render_headless", path, "3", out, "0", "0", "0", "2",
                          "3", "40", "30"], cwd=root, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    png = np.asarray(Image.open(out))
    s = G.load_gltf(path)
    cam = rt.Camera()
    _ffi.lib().rt_camera_frame_scene(s.handle, ctypes.byref(cam))
    with _renderer(s, width=40, height=30, spp=2, accumulate_frames=3, enable_denoise=False,
                   photon_count=50000) as r:
        r.update_camera(cam)
        r.render_frames(3)
        want = tonemap_cpu(r.read_accumulation())
    assert np.array_equal(png, want)
