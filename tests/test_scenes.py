"""Built-in scenes: area lights face the geometry they light.

The reference lights a point from an area-light triangle only when the
triangle's geometric normal cross(v1 - v0, v2 - v0) faces that point
(gather.wgsl:197-200), and emits photons only into that half-space
(photon_trace.wgsl:156-158). A light quad wound the other way lights nothing.
"""
import numpy as np
import pytest

import rustracer_amd as rt
from tests import helpers
from tests.oracle_ffi import OracleScene

# dark_room is the reference's Scene::dark_room (scene.rs:446-451), whose strip
# is wound towards +Y, away from the room: it stays as the reference builds it.
SYNTHETIC = [("cornell_box", 1), ("textured", 1), ("cornell36", 1), ("photon_plain", 1), ("multiview", 1),
             ("fullframe", 1)] + [("training", s) for s in range(1, 9)]


def _area_light_normals(scene):
    tris = scene.flatten_triangles()
    out = []
    for d in scene.light_descs():
        if d["kind"] != 2:
            continue
        t = tris[d["triangle_index"]]
        v0, v1, v2 = (np.asarray(t[k], np.float64) for k in ("v0", "v1", "v2"))
        out.append((np.cross(v1 - v0, v2 - v0), (v0 + v1 + v2) / 3.0))
    return out


@pytest.mark.parametrize("name,seed", SYNTHETIC)
def test_area_lights_face_scene_centre(name, seed):
    s = rt.Scene.builtin(name, seed)
    bmin, bmax = (np.asarray(v, np.float64) for v in s.bounds())
    centre = 0.5 * (bmin + bmax)
    lights = _area_light_normals(s)
    assert lights, f"{name} has no area light"
    for n, c in lights:
        assert np.dot(n, centre - c) > 0.0, f"{name}: area light at {c} faces away from the scene centre"


def test_dark_room_strip_is_the_reference_winding():
    # scene.rs:446-451: v0 (-1.5, 2.99, 0.15), v1 (1.5, 2.99, 0.15), v2 (1.5, 2.99, -0.15)
    # gives cross(v1 - v0, v2 - v0) = (0, 0.9, 0): +Y, as the reference has it.
    (n, _), _ = _area_light_normals(rt.Scene.builtin("dark_room"))
    assert n[1] > 0 and abs(n[0]) < 1e-6 and abs(n[2]) < 1e-6


def test_fullframe_nee_casts_shadow_rays():
    """At depth 1 with the photon map off, a pixel casts its camera ray plus one
    shadow ray whenever NEE picks an area light that faces the hit point. The
    environment light casts none, so a scene whose area lights face away would
    trace exactly one query per pixel."""
    s = rt.Scene.builtin("fullframe", 1)
    o = OracleScene.from_scene(s)
    w, h = 240, 135
    cfg = rt.RenderConfig(width=w, height=h, spp=1, enable_photon_map=False, max_depth=1)
    cam = rt.camera_to_uniform(rt.Scene.builtin_camera("fullframe", 1), w, h)
    _, _, _, st = o.gather(cam, helpers.image_params(cfg, 1), max_depth=1)
    per_px = st.ray_queries / (w * h)
    assert per_px > 1.25, per_px
