"""rustracer_amd — gfx950 (MI355X) backend of rustracer's rendering path.

Python mirror of the reference's crate API (rustracer-renderer/src/lib.rs,
rustracer-core/src/{scene,camera,material}.rs) over the C ABI in
include/rustracer_hip.h. Every call goes to librustracer_hip.so; there is no
CPU fallback, so a missing library or device raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import (BVH_NODE, HIT, LIGHT, MATERIAL, PHOTON, RAY, TRIANGLE, Camera, CameraUniform,
                   Config, GridParams, LightDesc, RtError, check, lib, ptr)

__all__ = ["RenderConfig", "Renderer", "Scene", "Camera", "RtError", "make_material", "camera",
           "TRIANGLE", "BVH_NODE", "MATERIAL", "LIGHT", "PHOTON", "RAY", "HIT"]


def RenderConfig(**overrides) -> Config:
    """RenderConfig::default() (lib.rs:43-64) with keyword overrides."""
    cfg = Config()
    check(lib().rt_config_default(C.byref(cfg)), "rt_config_default")
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise AttributeError(f"RenderConfig has no field {k!r}")
        if isinstance(v, bool):
            v = int(v)
        if k in ("light_emission", "point_light_emission"):
            setattr(cfg, k, (C.c_float * 3)(*v))
        else:
            setattr(cfg, k, v)
    return cfg


def make_material(kind: int, albedo=(0.8, 0.8, 0.8), roughness=0.5, metallic=0.0, ior=1.5,
                  emissive=(0.0, 0.0, 0.0)) -> np.ndarray:
    """One MaterialGpu record (material.rs:44-58) with no textures."""
    m = np.zeros(1, MATERIAL)
    m["albedo"][0] = (*albedo, 1.0)
    m["emissive"][0] = emissive
    m["roughness"], m["metallic"], m["ior"], m["kind"] = roughness, metallic, ior, kind
    for t in ("tex_albedo", "tex_mr", "tex_emissive", "tex_normal"):
        m[t] = _ffi.TEX_NONE
    return m


def project_marker(cam: Camera, width: int, height: int, position, radius: float) -> np.ndarray:
    """Renderer::project_marker (lib.rs:865-885): (x, y, radius in pixels, distance or 0)."""
    p = np.ascontiguousarray(position, np.float32)
    out = np.zeros(4, np.float32)
    check(lib().rt_project_marker(C.byref(cam), width, height, ptr(p), radius, ptr(out)), "rt_project_marker")
    return out


def camera(position, look_at, up=(0.0, 1.0, 0.0), fov_degrees=60.0, near_plane=0.1) -> Camera:
    c = Camera()
    check(lib().rt_camera_default(C.byref(c)), "rt_camera_default")
    c.position[:] = position
    c.look_at[:] = look_at
    c.up[:] = up
    c.fov_degrees = fov_degrees
    c.near_plane = near_plane
    return c


class Scene:
    """rustracer-core Scene (scene.rs:158-291) held by the native library."""

    def __init__(self, handle=None):
        if handle is None:
            h = C.c_void_p()
            check(lib().rt_scene_new(C.byref(h)), "rt_scene_new")
            handle = h
        self._h = handle

    @classmethod
    def builtin(cls, name: str, seed: int = 1) -> "Scene":
        h = C.c_void_p()
        check(lib().rt_scene_builtin(name.encode(), seed, C.byref(h)), f"rt_scene_builtin({name})")
        return cls(h)

    @staticmethod
    def builtin_camera(name: str, seed: int = 1, index: int = 0) -> Camera:
        c = Camera()
        check(lib().rt_scene_builtin_camera(name.encode(), seed, index, C.byref(c)), "rt_scene_builtin_camera")
        return c

    def __del__(self):
        try:
            if self._h:
                lib().rt_scene_free(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def add_material(self, material: np.ndarray) -> int:
        m = np.ascontiguousarray(material, MATERIAL).reshape(1)
        idx = C.c_uint32()
        check(lib().rt_scene_add_material(self._h, ptr(m), C.byref(idx)), "rt_scene_add_material")
        return idx.value

    def add_mesh(self, positions, indices, material_id: int, normals=None, uvs=None) -> None:
        p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        i = np.ascontiguousarray(indices, np.uint32).reshape(-1)
        n = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        u = None if uvs is None else np.ascontiguousarray(uvs, np.float32).reshape(-1, 2)
        check(lib().rt_scene_add_mesh(self._h, ptr(p), len(p), ptr(n), 0 if n is None else len(n),
                                      ptr(u), 0 if u is None else len(u), ptr(i), len(i), material_id),
              "rt_scene_add_mesh")

    def add_light(self, kind: int, vec_a=(0, 0, 0), vec_b=(0, 0, 0), color=(1, 1, 1), intensity=1.0,
                  triangle_index=0) -> None:
        d = LightDesc(kind, (C.c_float * 3)(*vec_a), (C.c_float * 3)(*vec_b), (C.c_float * 3)(*color),
                      intensity, triangle_index)
        check(lib().rt_scene_add_light(self._h, C.byref(d)), "rt_scene_add_light")

    def counts(self):
        n = [C.c_uint32() for _ in range(4)]
        check(lib().rt_scene_counts(self._h, *[C.byref(x) for x in n]), "rt_scene_counts")
        return dict(triangles=n[0].value, materials=n[1].value, lights=n[2].value, meshes=n[3].value)

    def bounds(self):
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        check(lib().rt_scene_bounds(self._h, ptr(lo), ptr(hi)), "rt_scene_bounds")
        return lo, hi

    def flatten_triangles(self) -> np.ndarray:
        out = np.zeros(self.counts()["triangles"], TRIANGLE)
        check(lib().rt_scene_flatten_triangles(self._h, ptr(out), len(out)), "rt_scene_flatten_triangles")
        return out

    def materials(self) -> np.ndarray:
        out = np.zeros(self.counts()["materials"], MATERIAL)
        check(lib().rt_scene_flatten_materials(self._h, ptr(out), len(out)), "rt_scene_flatten_materials")
        return out

    def light_descs(self) -> list:
        """The scene's lights as dicts (kind, vec_a, vec_b, color, intensity, triangle_index)."""
        n = C.c_uint32()
        check(lib().rt_scene_light_descs(self._h, None, 0, C.byref(n)), "rt_scene_light_descs")
        arr = (LightDesc * max(n.value, 1))()
        check(lib().rt_scene_light_descs(self._h, arr, n.value, C.byref(n)), "rt_scene_light_descs")
        return [dict(kind=d.kind, vec_a=tuple(d.vec_a), vec_b=tuple(d.vec_b), color=tuple(d.color),
                     intensity=d.intensity, triangle_index=d.triangle_index) for d in arr[:n.value]]

    def apply_descriptor(self, toml_text: str) -> None:
        """Apply a scene.toml descriptor (rustracer-core/src/config.rs:62-174)."""
        check(lib().rt_scene_apply_descriptor(self._h, toml_text.encode()), "rt_scene_apply_descriptor")

    def apply_descriptor_file(self, path: str) -> None:
        with open(path, encoding="utf-8") as f:
            self.apply_descriptor(f.read())

    def lights_gpu(self) -> np.ndarray:
        out = np.zeros(self.counts()["lights"], LIGHT)
        check(lib().rt_scene_lights_gpu(self._h, ptr(out), len(out)), "rt_scene_lights_gpu")
        return out

    def add_texture(self, role: int, rgba: np.ndarray) -> int:
        """Append an (h, w, 4) float RGBA texture to a role (0 albedo, 1 metallic-roughness,
        2 emissive, 3 normal); returns its layer index (the material's tex_* value)."""
        a = np.ascontiguousarray(rgba, np.float32)
        layer = C.c_uint32()
        check(lib().rt_scene_add_texture(self._h, role, a.shape[1], a.shape[0], ptr(a), C.byref(layer)),
              "rt_scene_add_texture")
        return layer.value

    def texture_array(self, role: int) -> np.ndarray:
        """The role's rgba8 array as uploaded: (layers, h, w, 4) uint8."""
        w, h, n = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().rt_scene_texture_array(self._h, role, None, 0, C.byref(w), C.byref(h), C.byref(n)),
              "rt_scene_texture_array")
        out = np.zeros((n.value, h.value, w.value, 4), np.uint8)
        check(lib().rt_scene_texture_array(self._h, role, ptr(out), out.size, C.byref(w), C.byref(h), C.byref(n)),
              "rt_scene_texture_array")
        return out

    def build_bvh(self):
        nn, nprim = C.c_uint32(), C.c_uint32()
        check(lib().rt_scene_build_bvh(self._h, None, 0, C.byref(nn), None, 0, C.byref(nprim)), "rt_scene_build_bvh")
        nodes = np.zeros(nn.value, BVH_NODE)
        prims = np.zeros(nprim.value, np.uint32)
        check(lib().rt_scene_build_bvh(self._h, ptr(nodes), len(nodes), C.byref(nn), ptr(prims), len(prims),
                                       C.byref(nprim)), "rt_scene_build_bvh")
        return nodes, prims

    def frame_camera(self) -> Camera:
        c = Camera()
        check(lib().rt_camera_frame_scene(self._h, C.byref(c)), "rt_camera_frame_scene")
        return c


def bvh_from_boxes(centroids, box_min, box_max):
    """BVHBuilder over explicit centroids and boxes (bvh.rs:127-151)."""
    c = np.ascontiguousarray(centroids, np.float32).reshape(-1, 3)
    lo = np.ascontiguousarray(box_min, np.float32).reshape(-1, 3)
    hi = np.ascontiguousarray(box_max, np.float32).reshape(-1, 3)
    n = len(c)
    nodes = np.zeros(max(2 * n, 1), BVH_NODE)
    prims = np.zeros(max(n, 1), np.uint32)
    nn = C.c_uint32()
    check(lib().rt_bvh_build_boxes(ptr(c), ptr(lo), ptr(hi), n, ptr(nodes), len(nodes), C.byref(nn), ptr(prims)),
          "rt_bvh_build_boxes")
    return nodes[:nn.value].copy(), prims[:n].copy()


def hbm_copy_probe(device: int = 0, nbytes: int = 1 << 30, iters: int = 20):
    """Measured HBM bandwidth of a plain copy kernel: (GB/s read+written, ms per copy)."""
    g, ms = C.c_double(), C.c_double()
    check(lib().rt_hbm_copy_probe(device, nbytes, iters, C.byref(g), C.byref(ms)), "rt_hbm_copy_probe")
    return g.value, ms.value


def camera_to_uniform(cam: Camera, width: int, height: int) -> CameraUniform:
    u = CameraUniform()
    check(lib().rt_camera_to_uniform(C.byref(cam), width, height, C.byref(u)), "rt_camera_to_uniform")
    return u


def grid_params_for(cfg: Config, bmin=None, bmax=None):
    gp = GridParams()
    radius = C.c_float()
    lo = None if bmin is None else np.ascontiguousarray(bmin, np.float32)
    hi = None if bmax is None else np.ascontiguousarray(bmax, np.float32)
    check(lib().rt_grid_params_for(C.byref(cfg), ptr(lo), ptr(hi), int(lo is not None), C.byref(gp),
                                   C.byref(radius)), "rt_grid_params_for")
    return gp, radius.value


class Renderer:
    """rustracer-renderer Renderer (lib.rs:189-1041) on one gfx950 device."""

    def __init__(self, config: Config, scene: Scene | None = None, device: int = 0, *, arrays=None):
        h = C.c_void_p()
        if arrays is None:
            check(lib().rt_renderer_create(C.byref(config), scene.handle, device, C.byref(h)), "rt_renderer_create")
        else:
            tris, nodes, prims, mats, lights, bounds = arrays[:6]
            tris = np.ascontiguousarray(tris, TRIANGLE)
            nodes = np.ascontiguousarray(nodes, BVH_NODE)
            prims = np.ascontiguousarray(prims, np.uint32)
            mats = np.ascontiguousarray(mats, MATERIAL)
            lights = np.ascontiguousarray(lights, LIGHT)
            lo = hi = None
            if bounds is not None:
                lo = np.ascontiguousarray(bounds[0], np.float32)
                hi = np.ascontiguousarray(bounds[1], np.float32)
            check(lib().rt_renderer_create_raw(C.byref(config), ptr(tris), len(tris), ptr(nodes), len(nodes),
                                               ptr(prims), len(prims), ptr(mats), len(mats), ptr(lights),
                                               len(lights), ptr(lo), ptr(hi), device, C.byref(h)),
                  "rt_renderer_create_raw")
        self._h = h
        cfg = self.config
        self.width, self.height = cfg.width, cfg.height
        if arrays is not None and len(arrays) > 6 and arrays[6] is not None:
            for role, t in enumerate(arrays[6]):
                if t is not None:
                    self.set_texture_array(role, t)

    def set_texture_array(self, role: int, texels: np.ndarray) -> None:
        """Upload a role's rgba8 array (layers, h, w, 4) uint8 (renderers made from raw arrays)."""
        t = np.ascontiguousarray(texels, np.uint8)
        check(lib().rt_renderer_set_texture_array(self._h, role, t.shape[2], t.shape[1], t.shape[0], ptr(t)),
              "rt_renderer_set_texture_array")

    def close(self):
        if getattr(self, "_h", None):
            lib().rt_renderer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- the reference's method set --
    def record_frame(self) -> None:
        check(lib().rt_render_frame(self._h), "rt_render_frame")

    def render_frames(self, n: int) -> None:
        check(lib().rt_render_frames(self._h, n), "rt_render_frames")

    def synchronize(self) -> None:
        check(lib().rt_synchronize(self._h), "rt_synchronize")

    def update_camera(self, cam: Camera) -> None:
        check(lib().rt_update_camera(self._h, C.byref(cam)), "rt_update_camera")

    def update_camera_uniform(self, u: CameraUniform) -> None:
        check(lib().rt_update_camera_uniform(self._h, C.byref(u)), "rt_update_camera_uniform")

    def update_lights(self, lights: np.ndarray) -> None:
        a = np.ascontiguousarray(lights, LIGHT)
        check(lib().rt_update_lights(self._h, ptr(a), len(a)), "rt_update_lights")

    def reset_accumulation(self) -> None:
        check(lib().rt_reset_accumulation(self._h), "rt_reset_accumulation")

    def reset_accumulation_soft(self) -> None:
        check(lib().rt_reset_accumulation_soft(self._h), "rt_reset_accumulation_soft")

    def rebuild_photons(self) -> None:
        check(lib().rt_rebuild_photons(self._h), "rt_rebuild_photons")

    def sample_count(self) -> int:
        return lib().rt_sample_count(self._h)

    def set_sample_count(self, v: int) -> None:
        check(lib().rt_set_sample_count(self._h, v), "rt_set_sample_count")

    @property
    def config(self) -> Config:
        c = Config()
        check(lib().rt_get_config(self._h, C.byref(c)), "rt_get_config")
        return c

    @config.setter
    def config(self, c: Config) -> None:
        check(lib().rt_set_config(self._h, C.byref(c)), "rt_set_config")

    def grid_params(self) -> GridParams:
        g = GridParams()
        check(lib().rt_get_grid_params(self._h, C.byref(g)), "rt_get_grid_params")
        return g

    # -- read-backs --
    def _px4(self, fn, name):
        out = np.zeros((self.height, self.width, 4), np.float32)
        check(fn(self._h, ptr(out), out.size), name)
        return out

    def read_accumulation(self):
        return self._px4(lib().rt_read_accumulation, "rt_read_accumulation")

    def read_frame(self):
        return self._px4(lib().rt_read_frame, "rt_read_frame")

    def read_moments(self):
        return self._px4(lib().rt_read_moments, "rt_read_moments")

    def read_gbuffer(self):
        return self._px4(lib().rt_read_gbuffer, "rt_read_gbuffer")

    def read_output_rgba8(self):
        out = np.zeros((self.height, self.width, 4), np.uint8)
        check(lib().rt_read_output_rgba8(self._h, ptr(out), out.size), "rt_read_output_rgba8")
        return out

    def read_present(self, surface_w: int, surface_h: int) -> np.ndarray:
        """The present pass (fullscreen.wgsl) into a surface_w x surface_h
        bgra8 surface, read back as (surface_h, surface_w, 4) bytes b, g, r, a."""
        out = np.zeros((surface_h, surface_w, 4), np.uint8)
        check(lib().rt_read_present(self._h, surface_w, surface_h, ptr(out), out.size), "rt_read_present")
        return out

    def present(self, d_surface: int, surface_w: int, surface_h: int, row_pitch: int) -> None:
        """The present pass into device memory at d_surface (row_pitch bytes per row)."""
        check(lib().rt_present(self._h, C.c_void_p(d_surface), surface_w, surface_h, row_pitch), "rt_present")

    def write_accumulation(self, rgba: np.ndarray) -> None:
        a = np.ascontiguousarray(rgba, np.float32)
        check(lib().rt_write_accumulation(self._h, ptr(a), a.size), "rt_write_accumulation")

    def read_photons(self) -> np.ndarray:
        n = C.c_uint32()
        check(lib().rt_read_photons(self._h, None, 0, C.byref(n)), "rt_read_photons")
        out = np.zeros(n.value, PHOTON)
        check(lib().rt_read_photons(self._h, ptr(out), len(out), C.byref(n)), "rt_read_photons")
        return out

    def read_sorted_photons(self) -> np.ndarray:
        n = C.c_uint32()
        check(lib().rt_read_sorted_photons(self._h, None, 0, C.byref(n)), "rt_read_sorted_photons")
        out = np.zeros(max(n.value, 1), PHOTON)
        check(lib().rt_read_sorted_photons(self._h, ptr(out), len(out), C.byref(n)), "rt_read_sorted_photons")
        return out[:n.value]

    def read_grid_meta(self) -> np.ndarray:
        n = C.c_uint32()
        check(lib().rt_read_grid_meta(self._h, None, 0, C.byref(n)), "rt_read_grid_meta")
        out = np.zeros((n.value, 2), np.uint32)
        check(lib().rt_read_grid_meta(self._h, ptr(out), out.size, C.byref(n)), "rt_read_grid_meta")
        return out

    # -- parity hooks --
    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(rays, RAY)
        out = np.zeros(len(r), HIT)
        check(lib().rt_trace_closest(self._h, ptr(r), len(r), ptr(out)), "rt_trace_closest")
        return out

    def trace_closest_device(self, d_rays: int, n: int, d_out: int) -> None:
        check(lib().rt_trace_closest_device(self._h, d_rays, n, d_out), "rt_trace_closest_device")

    def traversal_info(self) -> dict:
        """Active traversal (0 BVH2 nodes, 1 binary records, 2 nearest-first, 3 4-wide, 4 4-wide watertight)."""
        mode, depth, n4 = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().rt_traversal_info(self._h, C.byref(mode), C.byref(depth), C.byref(n4)),
              "rt_traversal_info")
        words = C.c_uint32()
        check(lib().rt_traversal_stack_words(self._h, C.byref(words)), "rt_traversal_stack_words")
        return {"mode": mode.value, "bvh_depth": depth.value, "records4": n4.value, "stack_words": words.value}

    def build_photon_grid_from(self, photons: np.ndarray) -> None:
        p = np.ascontiguousarray(photons, PHOTON)
        check(lib().rt_build_photon_grid_from(self._h, ptr(p), len(p)), "rt_build_photon_grid_from")

    # -- measurement --
    def pass_stats(self) -> dict:
        s = _ffi.PassStats()
        check(lib().rt_get_pass_stats(self._h, C.byref(s)), "rt_get_pass_stats")
        return {name: (s.ms[i], s.launches[i]) for i, name in enumerate(_ffi.PASS_NAMES)}

    def set_light_marker(self, position=None, radius: float = 0.0) -> None:
        """Draw the light gizmo at `position` (world units), or hide it with None."""
        p = None if position is None else np.ascontiguousarray(position, np.float32)
        check(lib().rt_set_light_marker(self._h, ptr(p), radius), "rt_set_light_marker")

    def set_light_glow(self, glow: float) -> None:
        check(lib().rt_set_light_glow(self._h, glow), "rt_set_light_glow")

    def frame_graph_captures(self) -> int:
        n = C.c_uint32()
        check(lib().rt_frame_graph_captures(self._h, C.byref(n)), "rt_frame_graph_captures")
        return n.value

    def reset_pass_stats(self) -> None:
        check(lib().rt_reset_pass_stats(self._h), "rt_reset_pass_stats")

    def count_gather_traffic(self) -> dict:
        t = _ffi.TraceCounters()
        check(lib().rt_count_gather_traffic(self._h, C.byref(t)), "rt_count_gather_traffic")
        phases = ("traversal", "shading", "photon_search", "shadow_traversal", "camera_loop")
        return dict(ray_queries=t.ray_queries, nodes_visited=t.nodes_visited, tris_tested=t.tris_tested,
                    photon_reads=t.photon_reads,
                    active_lanes={p: round(t.lane_steps[k] / max(t.wave_steps[k], 1), 2) for k, p in enumerate(phases)},
                    lane_steps={p: t.lane_steps[k] for k, p in enumerate(phases)},
                    wave_steps={p: t.wave_steps[k] for k, p in enumerate(phases)},
                    wave_cycles={p: t.wave_cycles[k] for k, p in enumerate(phases)})

    # -- many views in one persistent launch --
    def set_views(self, views) -> None:
        """views: sequence of (Camera, width, height, spp, frame)."""
        arr = (_ffi.View * len(views))()
        self._views = []
        for k, (cam, w, h, spp, frame) in enumerate(views):
            arr[k].camera = camera_to_uniform(cam, w, h)
            arr[k].width, arr[k].height, arr[k].spp, arr[k].frame = w, h, spp, frame
            self._views.append((w, h))
        check(lib().rt_batch_set_views(self._h, arr, len(views)), "rt_batch_set_views")

    def render_batch(self, tile_w=None, tile_h=8, waves_per_cu=16, tiles_per_fetch=1, lanes=1) -> None:
        """One launch over all views; tile_w None keeps the current tuning."""
        t = None
        if tile_w is not None:
            t = C.byref(_ffi.BatchTuning(tile_w, tile_h, waves_per_cu, tiles_per_fetch, lanes))
        check(lib().rt_render_batch(self._h, t), "rt_render_batch")

    def batch_autotune(self, repeats: int = 1):
        best = _ffi.BatchTuning()
        ms = C.c_double()
        check(lib().rt_batch_autotune(self._h, repeats, C.byref(best), C.byref(ms)), "rt_batch_autotune")
        return best.as_dict(), ms.value

    def batch_totals(self):
        px, smp = C.c_uint64(), C.c_uint64()
        check(lib().rt_batch_totals(self._h, C.byref(px), C.byref(smp)), "rt_batch_totals")
        return px.value, smp.value

    def read_batch_view(self, k: int):
        w, h = self._views[k]
        out = [np.zeros((h, w, 4), np.float32) for _ in range(3)]
        check(lib().rt_read_batch_view(self._h, k, ptr(out[0]), ptr(out[1]), ptr(out[2]), out[0].size),
              "rt_read_batch_view")
        return tuple(out)

    def accum64_reset(self) -> None:
        check(lib().rt_accum64_reset(self._h), "rt_accum64_reset")

    def accum64_add(self) -> None:
        check(lib().rt_accum64_add_accumulation(self._h), "rt_accum64_add_accumulation")

    def accum64_mean(self):
        out = np.zeros((self.height, self.width, 4), np.float64)
        n = C.c_uint64()
        check(lib().rt_accum64_read_mean(self._h, ptr(out), out.size, C.byref(n)), "rt_accum64_read_mean")
        return out, n.value

    def accum64_relmse_display(self, eps: float = 1e-2) -> float:
        """relMSE of the displayed image (the learned denoiser's output when attached)."""
        v = C.c_double()
        check(lib().rt_accum64_relmse_display(self._h, eps, C.byref(v)), "rt_accum64_relmse_display")
        return v.value

    def set_denoiser(self, unet) -> None:
        """Attach a learned denoiser (rustracer_amd.denoiser.UNet) or detach with None."""
        check(lib().rt_renderer_set_denoiser(self._h, None if unet is None else unet.handle),
              "rt_renderer_set_denoiser")
        self._unet = unet  # keep it alive while attached

    def read_display(self):
        return self._px4(lib().rt_read_display, "rt_read_display")

    def read_albedo(self):
        return self._px4(lib().rt_read_albedo, "rt_read_albedo")

    def accum64_relmse(self, eps: float = 1e-2, d_image: int | None = None) -> float:
        """relMSE of the accumulation (or a device float4 image) against the FP64 mean."""
        v = C.c_double()
        check(lib().rt_accum64_relmse(self._h, d_image, eps, C.byref(v)), "rt_accum64_relmse")
        return v.value
