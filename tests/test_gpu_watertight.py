"""The opt-in watertight mode (RT_FLAG_WATERTIGHT, traversal mode 4): the
watertight ray/triangle test behind conservative box tests. Its closest hits
equal, bit for bit, a watertight test of every triangle (the oracle's brute
force); its differences from the reference's Moller-Trumbore are counted and
bounded per ray set; and rays aimed at the shared edges of a closed mesh from
inside it never escape (Moller-Trumbore's escapes are counted)."""
import numpy as np
import pytest

import rustracer_amd as rt
from rustracer_amd import _ffi

from .helpers import camera_rays, face_rays, icosphere, random_rays
from .oracle_ffi import OracleScene

pytestmark = pytest.mark.gpu

# Differences from Moller-Trumbore (the reference's test and traversal), per
# ray set. Camera and random rays meet edges only by chance: hit versus miss may
# differ on at most HIT_MISS_BOUND of them, and a different triangle is allowed
# only as a near-tie (distances within TIE_ABS of the scene's diagonal:
# coincident faces, shared edges; the watertight t is good to a few ulps of the
# vertex distances, not of t). Face rays lie in the plane of a BVH box face and
# so run along the edges on it: the reference's traversal culls a box whose min
# face holds the ray, and Moller-Trumbore's rounding decides edge contacts,
# while the watertight mode counts an edge contact as a hit. There every ray
# that the watertight mode hits nearer, or alone, must touch its triangle on an
# edge (a barycentric within EDGE_BARY of 0), and Moller-Trumbore may hit alone
# on at most MT_ONLY_BOUND of the rays.
HIT_MISS_BOUND = 1e-3
TIE_ABS = 1e-5
EDGE_BARY = 1e-4
MT_ONLY_BOUND = 0.01


def _min_barycentric(tris, rays, t, prim):
    """float64 smallest barycentric of the point o + t d on triangle prim."""
    v0 = tris["v0"][prim].astype(np.float64)
    e1 = tris["v1"][prim].astype(np.float64) - v0
    e2 = tris["v2"][prim].astype(np.float64) - v0
    p = rays["origin"].astype(np.float64) + rays["dir"].astype(np.float64) * t[:, None].astype(np.float64) - v0
    n = np.cross(e1, e2)
    nn = np.einsum("ij,ij->i", n, n)
    b1 = np.einsum("ij,ij->i", np.cross(p, e2), n) / nn
    b2 = np.einsum("ij,ij->i", np.cross(e1, p), n) / nn
    return np.minimum(np.minimum(b1, b2), 1 - b1 - b2)


def _renderer(scene, flags):
    return rt.Renderer(rt.RenderConfig(width=8, height=8, flags=flags), scene, 0)


def _check_against_brute_force(orc, rays, wt):
    ref = orc.brute_watertight(rays)
    np.testing.assert_array_equal(wt["t"], ref["t"])
    # the same triangle, or another one at exactly the same distance (the
    # brute force takes the lowest index, traversal the first one it meets)
    d = (wt["prim_id"] != ref["prim_id"]) & (ref["t"] < 1e38)
    if d.any():
        np.testing.assert_array_equal(orc.watertight_t(rays[d], wt["prim_id"][d]), wt["t"][d])
    return int(d.sum())


@pytest.mark.parametrize("scene_name", ["cornell_box", "photon_plain", "multiview"])
def test_watertight_against_brute_force_and_moller_trumbore(scene_name):
    scene = rt.Scene.builtin(scene_name, 3)
    lo, hi = scene.bounds()
    nodes, _ = scene.build_bvh()
    sets = {"camera": camera_rays(rt.Scene.builtin_camera(scene_name, 3), 96, 64, 1),
            "random": random_rays(lo, hi, 20000, 7), "face": face_rays(nodes, 5000, 11)}
    rays = np.concatenate(list(sets.values()))
    diag = float(np.linalg.norm(np.asarray(hi, np.float64) - np.asarray(lo, np.float64)))
    print(f"watertight {scene_name} rays={len(rays)}", flush=True)
    with _renderer(scene, 0) as r:
        assert r.traversal_info()["mode"] == 3
        mt = r.trace_closest(rays)
    with _renderer(scene, _ffi.FLAG_WATERTIGHT) as r:
        assert r.traversal_info()["mode"] == 4
        wt = r.trace_closest(rays)
    ties = _check_against_brute_force(OracleScene.from_scene(scene), rays, wt)
    tris = scene.flatten_triangles()
    print(f"  equal to the brute force on every ray ({ties} exact ties to another triangle)", flush=True)
    start = 0
    for name, rs in sets.items():
        m, w = mt[start:start + len(rs)], wt[start:start + len(rs)]
        start += len(rs)
        hm, hw = m["t"] < 1e38, w["t"] < 1e38
        hit_miss = int((hm != hw).sum())
        other = int((hm & hw & (m["prim_id"] != w["prim_id"])).sum())
        print(f"  {name}: {len(rs)} rays, {int(hm.sum())} hits; Moller-Trumbore differs: hit/miss {hit_miss}, "
              f"triangle {other}", flush=True)
        assert hm.sum() > len(rs) // 20
        d = hm & hw & (m["prim_id"] != w["prim_id"])
        same = hm & hw & (m["prim_id"] == w["prim_id"])
        gap = np.abs(w["t"][same] - m["t"][same])
        # the same triangle: distances agree to rounding of the scene's size,
        # and typically to a few ulps of t
        assert gap.max() <= TIE_ABS * diag, gap.max()
        assert np.median(gap / np.maximum(m["t"][same], 1e-6)) < 1e-6
        if name != "face":
            assert hit_miss / len(rs) <= HIT_MISS_BOUND
            assert (np.abs(w["t"][d] - m["t"][d]) <= TIE_ABS * diag).all(), np.abs(w["t"][d] - m["t"][d]).max()
            continue
        nearer = (hw & ~hm) | (d & (w["t"] < m["t"] - TIE_ABS * diag))
        if nearer.any():
            b = _min_barycentric(tris, rs[nearer], w["t"][nearer], w["prim_id"][nearer])
            print(f"    watertight nearer or alone on {int(nearer.sum())} rays; largest min barycentric {b.max():.2e}",
                  flush=True)
            assert np.abs(b).max() < EDGE_BARY
        farther = d & (w["t"] > m["t"] + TIE_ABS * diag)
        mt_only = int((hm & ~hw).sum()) + int(farther.sum())
        assert mt_only / len(rs) <= MT_ONLY_BOUND, mt_only


def test_no_ray_escapes_a_closed_mesh_through_its_edges():
    V, Fc = icosphere(4, 5)
    s = rt.Scene()
    m = s.add_material(rt.make_material(0))
    s.add_mesh(V, Fc.ravel(), m)
    edges = np.array(sorted({(min(a, b), max(a, b)) for tri in Fc for a, b in ((tri[0], tri[1]), (tri[1], tri[2]),
                                                                             (tri[2], tri[0]))}))
    rng = np.random.default_rng(9)
    k = rng.integers(0, len(edges), 200_000)
    w = rng.random(k.shape)[:, None].astype(np.float32)
    p = V[edges[k, 0]] * (np.float32(1) - w) + V[edges[k, 1]] * w  # points on shared edges (float32)
    o = (0.05 * rng.standard_normal((len(k), 3))).astype(np.float32)  # origins inside the mesh
    d = (p - o).astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(len(k), _ffi.RAY)
    rays["origin"] = o
    rays["dir"] = d.astype(np.float32)
    print(f"closed icosphere: {len(Fc)} triangles, {len(rays)} rays at shared edges", flush=True)
    with _renderer(s, _ffi.FLAG_WATERTIGHT) as r:
        assert r.traversal_info()["mode"] == 4
        wt = r.trace_closest(rays)
    with _renderer(s, 0) as r:
        mt = r.trace_closest(rays)
    escapes_mt = int((mt["t"] >= 1e38).sum())
    print(f"  escapes: watertight 0 expected, got {int((wt['t'] >= 1e38).sum())}; Moller-Trumbore {escapes_mt}",
          flush=True)
    assert (wt["t"] < 1e38).all()
    _check_against_brute_force(OracleScene.from_scene(s), rays, wt)


def test_watertight_frames_render():
    scene = rt.Scene.builtin("photon_plain", 2)
    cfg = rt.RenderConfig(width=48, height=32, spp=2, max_depth=6, enable_photon_map=True, photon_count=20000,
                          flags=_ffi.FLAG_WATERTIGHT)
    with rt.Renderer(cfg, scene, 0) as r:
        assert r.traversal_info()["mode"] == 4
        r.update_camera(rt.Scene.builtin_camera("photon_plain", 2))
        r.rebuild_photons()
        r.render_frames(3)
        frame, gb = r.read_frame(), r.read_gbuffer()
    assert np.isfinite(frame).all() and np.isfinite(gb).all()
    assert (frame[..., :3] > 0).mean() > 0.5
