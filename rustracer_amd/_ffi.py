"""ctypes binding of librustracer_hip.so (include/rustracer_hip.h, include/rt_types.h).

The library is built in-tree by ``rustracer_amd/Makefile``; loading fails loudly
when it is missing -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librustracer_hip.so")

# ---------------------------------------------------------------- record dtypes
f3 = (np.float32, 3)
TRIANGLE = np.dtype([("v0", f3), ("_pad0", np.float32), ("v1", f3), ("_pad1", np.float32),
                     ("v2", f3), ("_pad2", np.float32), ("n0", f3), ("_pad3", np.float32),
                     ("n1", f3), ("_pad4", np.float32), ("n2", f3), ("_pad5", np.float32),
                     ("uv0", (np.float32, 2)), ("uv1", (np.float32, 2)), ("uv2", (np.float32, 2)),
                     ("material_id", np.uint32), ("_pad6", np.float32)])
BVH_NODE = np.dtype([("bbox_min", f3), ("left_first", np.uint32), ("bbox_max", f3),
                     ("right_child", np.uint32), ("prim_count", np.uint32), ("_pad", f3)])
MATERIAL = np.dtype([("albedo", (np.float32, 4)), ("emissive", f3), ("roughness", np.float32),
                     ("metallic", np.float32), ("ior", np.float32), ("kind", np.uint32),
                     ("flags", np.uint32), ("tex_albedo", np.uint32), ("tex_mr", np.uint32),
                     ("tex_emissive", np.uint32), ("tex_normal", np.uint32)])
LIGHT = np.dtype([("data", (np.float32, 4)), ("color", f3), ("intensity", np.float32),
                  ("kind", np.uint32), ("_pad_a", f3), ("_pad_b", (np.float32, 4))])
PHOTON = np.dtype([("position", f3), ("_pad0", np.float32), ("power", f3), ("_pad1", np.float32),
                   ("incident", f3), ("_pad2", np.float32), ("normal", f3), ("_pad3", np.float32)])
RAY = np.dtype([("origin", f3), ("dir", f3)])
HIT = np.dtype([("t", np.float32), ("prim_id", np.uint32)])
for _dt, _size in ((TRIANGLE, 128), (BVH_NODE, 48), (MATERIAL, 64), (LIGHT, 64), (PHOTON, 64),
                   (RAY, 24), (HIT, 8)):
    assert _dt.itemsize == _size, (_dt, _size)

TEX_NONE = 0xFFFFFFFF
MAT_LAMBERTIAN, MAT_METAL, MAT_DIELECTRIC, MAT_EMISSIVE = 0, 1, 2, 3
LIGHT_POINT, LIGHT_DIRECTIONAL, LIGHT_AREA, LIGHT_ENVIRONMENT = 0, 1, 2, 3
GRID_MAX_CELLS = 26 * 26 * 26

RT_OK, RT_ERR_INVALID, RT_ERR_UNSUPPORTED, RT_ERR_HIP, RT_ERR_NO_DEVICE, RT_ERR_CAPACITY = 0, -1, -2, -3, -4, -5
FLAG_PROFILE, FLAG_CLOSEST_SHADOW, FLAG_REFERENCE_TRAVERSAL, FLAG_NEAREST_FIRST = 1, 2, 4, 8
FLAG_BINARY_RECORDS = 16
FLAG_PATH_QUEUES = 32
FLAG_TILE_LAUNCH = 64
FLAG_GRAPH = 128
FLAG_MOMENT_HISTORY = 256
FLAG_WATERTIGHT = 512
PASS_NAMES = ("gather", "denoise", "composite_tonemap", "photon_trace", "grid_build", "trace_rays",
              "accum64", "batch", "network", "present")


class Config(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("photon_count", C.c_uint32),
                ("spp", C.c_uint32), ("exposure", C.c_float), ("gamma", C.c_float),
                ("light_emission", C.c_float * 3), ("point_light_emission", C.c_float * 3),
                ("accumulate_frames", C.c_uint32), ("enable_photon_map", C.c_uint32),
                ("enable_denoise", C.c_uint32), ("photon_debug", C.c_uint32),
                ("debug_view", C.c_uint32), ("photon_scale", C.c_float),
                ("photon_radius", C.c_float), ("ambient", C.c_float), ("max_depth", C.c_uint32),
                ("grid_cell_size", C.c_float), ("denoise_passes", C.c_uint32), ("flags", C.c_uint32)]


class CameraUniform(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("_pad0", C.c_float), ("forward", C.c_float * 3),
                ("_pad1", C.c_float), ("right", C.c_float * 3), ("_pad2", C.c_float),
                ("up", C.c_float * 3), ("_pad3", C.c_float), ("inv_half_res", C.c_float * 2),
                ("near_plane", C.c_float), ("fov_scale", C.c_float), ("aperture", C.c_float),
                ("focus_dist", C.c_float), ("_pad_end", C.c_float * 2)]


class ImageParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("frame", C.c_uint32), ("light_emission", C.c_float * 3), ("ambient", C.c_float),
                ("accumulate_alpha", C.c_float), ("use_photons", C.c_uint32),
                ("photon_count", C.c_uint32), ("photon_radius", C.c_float),
                ("photon_scale", C.c_float), ("photon_debug", C.c_uint32),
                ("point_light_emission", C.c_float * 3), ("_pad", C.c_uint32 * 3)]


class GridParams(C.Structure):
    _fields_ = [("room_min_x", C.c_float), ("room_min_y", C.c_float), ("room_min_z", C.c_float),
                ("cell_size", C.c_float), ("grid_x", C.c_uint32), ("grid_y", C.c_uint32),
                ("grid_z", C.c_uint32), ("num_cells", C.c_uint32), ("photon_slots", C.c_uint32),
                ("_pad", C.c_uint32 * 3)]


class DenoiseParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("step_size", C.c_uint32),
                ("lambda_", C.c_float), ("phi_normal", C.c_float), ("phi_depth", C.c_float),
                ("alpha", C.c_float), ("_pad", C.c_uint32)]


class ToneMapUniform(C.Structure):
    _fields_ = [("exposure", C.c_float), ("gamma", C.c_float), ("width", C.c_uint32),
                ("glow", C.c_float), ("marker", C.c_float * 4)]


class LightDesc(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("vec_a", C.c_float * 3), ("vec_b", C.c_float * 3),
                ("color", C.c_float * 3), ("intensity", C.c_float), ("triangle_index", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("look_at", C.c_float * 3), ("up", C.c_float * 3),
                ("fov_degrees", C.c_float), ("near_plane", C.c_float), ("aperture", C.c_float),
                ("focus_dist", C.c_float)]


class PassStats(C.Structure):
    _fields_ = [("ms", C.c_double * 12), ("launches", C.c_uint64 * 12)]


class View(C.Structure):
    _fields_ = [("camera", CameraUniform), ("width", C.c_uint32), ("height", C.c_uint32),
                ("spp", C.c_uint32), ("frame", C.c_uint32)]


class BatchTuning(C.Structure):
    _fields_ = [("tile_w", C.c_uint32), ("tile_h", C.c_uint32), ("waves_per_cu", C.c_uint32),
                ("tiles_per_fetch", C.c_uint32), ("lanes", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


COUNT_PHASES = 5  # RT_COUNT_PHASES


class TraceCounters(C.Structure):
    _fields_ = [("ray_queries", C.c_uint64), ("nodes_visited", C.c_uint64),
                ("tris_tested", C.c_uint64), ("photon_reads", C.c_uint64),
                ("lane_steps", C.c_uint64 * COUNT_PHASES), ("wave_steps", C.c_uint64 * COUNT_PHASES),
                ("wave_cycles", C.c_uint64 * COUNT_PHASES)]


for _s, _size in ((Config, 96), (CameraUniform, 96), (ImageParams, 80), (GridParams, 48),
                  (DenoiseParams, 32), (ToneMapUniform, 32)):
    assert C.sizeof(_s) == _size, (_s, C.sizeof(_s))

_vp = C.c_void_p
_u32 = C.c_uint32
_i = C.c_int
_P = C.POINTER

# name -> (restype, argtypes); every symbol of include/rustracer_hip.h
SIGNATURES = {
    "rt_last_error": (C.c_char_p, []),
    "rt_version": (C.c_char_p, []),
    "rt_config_default": (_i, [_P(Config)]),
    "rt_scene_new": (_i, [_P(_vp)]),
    "rt_scene_free": (None, [_vp]),
    "rt_scene_add_material": (_i, [_vp, _vp, _P(_u32)]),
    "rt_scene_add_mesh": (_i, [_vp, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32, _u32]),
    "rt_scene_add_light": (_i, [_vp, _P(LightDesc)]),
    "rt_scene_clear_lights": (_i, [_vp]),
    "rt_scene_builtin": (_i, [C.c_char_p, C.c_uint64, _P(_vp)]),
    "rt_scene_builtin_camera": (_i, [C.c_char_p, C.c_uint64, _u32, _P(Camera)]),
    "rt_scene_counts": (_i, [_vp, _P(_u32), _P(_u32), _P(_u32), _P(_u32)]),
    "rt_scene_bounds": (_i, [_vp, _vp, _vp]),
    "rt_scene_flatten_triangles": (_i, [_vp, _vp, _u32]),
    "rt_scene_flatten_materials": (_i, [_vp, _vp, _u32]),
    "rt_scene_lights_gpu": (_i, [_vp, _vp, _u32]),
    "rt_scene_light_descs": (_i, [_vp, _vp, C.c_uint32, _vp]),
    "rt_tonemap_cpu": (_i, [_vp, C.c_size_t, _vp]),
    "rt_srgb_to_linear": (_i, [_vp, C.c_size_t]),
    "rt_scene_apply_descriptor": (_i, [_vp, C.c_char_p]),
    "rt_light_to_gpu": (_i, [_P(LightDesc), _vp]),
    "rt_scene_build_bvh": (_i, [_vp, _vp, _u32, _P(_u32), _vp, _u32, _P(_u32)]),
    "rt_bvh_build_boxes": (_i, [_vp, _vp, _vp, _u32, _vp, _u32, _P(_u32), _vp]),
    "rt_camera_default": (_i, [_P(Camera)]),
    "rt_camera_to_uniform": (_i, [_P(Camera), _u32, _u32, _P(CameraUniform)]),
    "rt_camera_frame_scene": (_i, [_vp, _P(Camera)]),
    "rt_camera_orbit": (_i, [_P(Camera), C.c_float, C.c_float]),
    "rt_camera_zoom": (_i, [_P(Camera), C.c_float]),
    "rt_camera_pan": (_i, [_P(Camera), C.c_float, C.c_float]),
    "rt_grid_params_for": (_i, [_P(Config), _vp, _vp, _i, _P(GridParams), _P(C.c_float)]),
    "rt_renderer_create": (_i, [_P(Config), _vp, _i, _P(_vp)]),
    "rt_renderer_create_raw": (_i, [_P(Config), _vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32, _vp,
                                    _u32, _vp, _vp, _i, _P(_vp)]),
    "rt_renderer_destroy": (None, [_vp]),
    "rt_render_frame": (_i, [_vp]),
    "rt_render_frames": (_i, [_vp, _u32]),
    "rt_synchronize": (_i, [_vp]),
    "rt_update_camera": (_i, [_vp, _P(Camera)]),
    "rt_update_camera_uniform": (_i, [_vp, _P(CameraUniform)]),
    "rt_update_lights": (_i, [_vp, _vp, _u32]),
    "rt_reset_accumulation": (_i, [_vp]),
    "rt_reset_accumulation_soft": (_i, [_vp]),
    "rt_rebuild_photons": (_i, [_vp]),
    "rt_sample_count": (_u32, [_vp]),
    "rt_set_sample_count": (_i, [_vp, _u32]),
    "rt_get_config": (_i, [_vp, _P(Config)]),
    "rt_set_config": (_i, [_vp, _P(Config)]),
    "rt_get_grid_params": (_i, [_vp, _P(GridParams)]),
    "rt_read_accumulation": (_i, [_vp, _vp, C.c_size_t]),
    "rt_read_frame": (_i, [_vp, _vp, C.c_size_t]),
    "rt_read_moments": (_i, [_vp, _vp, C.c_size_t]),
    "rt_read_gbuffer": (_i, [_vp, _vp, C.c_size_t]),
    "rt_read_output_rgba8": (_i, [_vp, _vp, C.c_size_t]),
    "rt_write_accumulation": (_i, [_vp, _vp, C.c_size_t]),
    "rt_present": (_i, [_vp, _vp, _u32, _u32, C.c_size_t]),
    "rt_read_present": (_i, [_vp, _u32, _u32, _vp, C.c_size_t]),
    "rt_read_photons": (_i, [_vp, _vp, C.c_size_t, _P(_u32)]),
    "rt_read_sorted_photons": (_i, [_vp, _vp, C.c_size_t, _P(_u32)]),
    "rt_read_grid_meta": (_i, [_vp, _vp, C.c_size_t, _P(_u32)]),
    "rt_trace_closest": (_i, [_vp, _vp, _u32, _vp]),
    "rt_trace_closest_device": (_i, [_vp, _vp, _u32, _vp]),
    "rt_traversal_stack_words": (_i, [_vp, _vp]),
    "rt_traversal_info": (_i, [_vp, _vp, _vp, _vp]),
    "rt_set_light_marker": (_i, [_vp, _vp, C.c_float]),
    "rt_set_light_glow": (_i, [_vp, C.c_float]),
    "rt_project_marker": (_i, [_P(Camera), C.c_uint32, C.c_uint32, _vp, C.c_float, _vp]),
    "rt_frame_graph_captures": (_i, [_vp, _vp]),
    "rt_renderer_set_texture_array": (_i, [_vp, _u32, _u32, _u32, _u32, _vp]),
    "rt_scene_add_texture": (_i, [_vp, _u32, _u32, _u32, _vp, _P(_u32)]),
    "rt_scene_texture_array": (_i, [_vp, _u32, _vp, C.c_size_t, _P(_u32), _P(_u32), _P(_u32)]),
    "rt_build_photon_grid_from": (_i, [_vp, _vp, _u32]),
    "rt_photon_grid_build_noalloc": (_i, [_vp, _u32, _P(GridParams), _vp, _u32, _vp, _vp, _vp, _vp]),
    "rt_get_pass_stats": (_i, [_vp, _P(PassStats)]),
    "rt_reset_pass_stats": (_i, [_vp]),
    "rt_count_gather_traffic": (_i, [_vp, _P(TraceCounters)]),
    "rt_hbm_copy_probe": (_i, [_i, C.c_size_t, _u32, _P(C.c_double), _P(C.c_double)]),
    "rt_batch_set_views": (_i, [_vp, _vp, _u32]),
    "rt_render_batch": (_i, [_vp, _vp]),
    "rt_batch_autotune": (_i, [_vp, _u32, _vp, _P(C.c_double)]),
    "rt_batch_get_tuning": (_i, [_vp, _vp]),
    "rt_batch_totals": (_i, [_vp, _P(C.c_uint64), _P(C.c_uint64)]),
    "rt_read_batch_view": (_i, [_vp, _u32, _vp, _vp, _vp, C.c_size_t]),
    "rt_accum64_reset": (_i, [_vp]),
    "rt_accum64_add_accumulation": (_i, [_vp]),
    "rt_accum64_read_mean": (_i, [_vp, _vp, C.c_size_t, _P(C.c_uint64)]),
    "rt_accum64_relmse": (_i, [_vp, _vp, C.c_float, _P(C.c_double)]),
    "rt_accum64_relmse_display": (_i, [_vp, C.c_float, _P(C.c_double)]),
    "rt_unet_create": (_i, [C.c_int, C.c_uint64, _P(_vp)]),
    "rt_unet_destroy": (_i, [_vp]),
    "rt_unet_param_count": (_i, [_vp, _P(C.c_size_t)]),
    "rt_unet_get_params": (_i, [_vp, _vp, C.c_size_t]),
    "rt_unet_set_params": (_i, [_vp, _vp, C.c_size_t]),
    "rt_unet_dataset_reset": (_i, [_vp, C.c_uint32, C.c_uint32, C.c_uint32]),
    "rt_unet_dataset_set": (_i, [_vp, C.c_uint32, C.c_int, _vp]),
    "rt_unet_train": (_i, [_vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, _vp]),
    "rt_unet_eval": (_i, [_vp, C.c_uint32, C.c_uint32, C.c_uint32, _vp, _vp, _vp, _vp, _vp, _P(C.c_double)]),
    "rt_unet_conv_choices": (_i, [_vp, _vp, C.c_size_t, _P(C.c_size_t)]),
    "rt_unet_set_variant_mode": (_i, [_vp, C.c_int]),
    "rt_renderer_set_denoiser": (_i, [_vp, _vp]),
    "rt_read_display": (_i, [_vp, _vp, C.c_size_t]),
    "rt_read_albedo": (_i, [_vp, _vp, C.c_size_t]),
}


class RtError(RuntimeError):
    def __init__(self, code: int, where: str, message: str):
        super().__init__(f"{where} failed ({code}): {message}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    """Load the in-tree library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `make -C rustracer_amd` "
                              "(or __graft_entry__.build()) first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code: int, where: str) -> None:
    if code != RT_OK:
        msg = lib().rt_last_error().decode("utf-8", "replace")
        raise RtError(code, where, msg)


def ptr(a: np.ndarray | None) -> int | None:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the C ABI must be contiguous"
    return a.ctypes.data
