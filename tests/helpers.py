"""Shared helpers: reference-side parameter records and the oracle frame pipeline."""
from __future__ import annotations

import numpy as np

import rustracer_amd as rt
from rustracer_amd import _ffi

from . import oracle_ffi as orc


def image_params(cfg, frame: int, photon_radius: float | None = None, alpha: float = 1.0) -> _ffi.ImageParams:
    """ImageParams as Renderer::record_frame writes them (lib.rs:685-703)."""
    ip = _ffi.ImageParams()
    ip.width, ip.height, ip.spp, ip.frame = cfg.width, cfg.height, cfg.spp, frame
    ip.light_emission[:] = tuple(cfg.light_emission)
    ip.ambient = cfg.ambient
    ip.accumulate_alpha = alpha
    ip.use_photons = 1 if cfg.enable_photon_map else 0
    ip.photon_count = cfg.photon_count
    ip.photon_radius = cfg.photon_radius if photon_radius is None else photon_radius
    ip.photon_scale = cfg.photon_scale
    ip.photon_debug = 1 if cfg.photon_debug else 0
    ip.point_light_emission[:] = tuple(cfg.point_light_emission)
    ip._pad[0] = cfg.debug_view
    return ip


def alpha_for(accumulate_frames: int, frames_since_reset: int) -> float:
    """lib.rs:679-684 in float32."""
    if accumulate_frames == 0:
        return 1.0
    return float(np.float32(1.0) / np.float32(min(frames_since_reset, accumulate_frames)))


def oracle_frames(oscene: orc.OracleScene, cfg, cam_uniform, n_frames: int, photon_map=None, grid_radius=None,
                  threads: int = 8):
    """Renderer::record_frame x n_frames on the oracle: gather, a-trous passes,
    composite, tonemap. Returns (accumulation, rgba8, last frame, moments, gbuffer)."""
    h, w = cfg.height, cfg.width
    acc = np.zeros((h, w, 4), np.float32)
    rgba = None
    frame = mom = gb = None
    for k in range(1, n_frames + 1):
        alpha = alpha_for(cfg.accumulate_frames, k)
        ip = image_params(cfg, k, grid_radius, alpha)
        frame, mom, gb, _ = oscene.gather(cam_uniform, ip, max_depth=cfg.max_depth, photon_map=photon_map,
                                          threads=threads)
        cur = frame
        if cfg.enable_denoise:
            for it in range(cfg.denoise_passes):
                p = _ffi.DenoiseParams(w, h, 1 << it, 3.0, 12.0, 0.05, alpha, 0)
                cur = orc.denoise_pass(cur, gb, mom, p)
        acc = orc.composite(cur, acc, alpha)
        tm = _ffi.ToneMapUniform(cfg.exposure, cfg.gamma, w, 1.0, (0.0, 0.0, 0.0, 0.0))
        rgba = orc.tonemap(acc, gb, tm)
    return acc, rgba, frame, mom, gb


def camera_rays(cam, w: int, h: int, jitter_seed: int = 0) -> np.ndarray:
    """Primary rays of a pinhole camera through jittered pixel positions."""
    u = rt.camera_to_uniform(cam, w, h)
    rng = np.random.default_rng(jitter_seed)
    ys, xs = np.mgrid[0:h, 0:w]
    jx = rng.uniform(-0.5, 0.5, xs.shape).astype(np.float32)
    jy = rng.uniform(-0.5, 0.5, ys.shape).astype(np.float32)
    ux = (xs.astype(np.float32) + jx) * np.float32(u.inv_half_res[0]) - np.float32(1)
    uy = np.float32(1) - (ys.astype(np.float32) + jy) * np.float32(u.inv_half_res[1])
    fwd, right, up = (np.array(v, np.float32) for v in (u.forward, u.right, u.up))
    d = fwd[None, None] + right[None, None] * ux[..., None] + up[None, None] * uy[..., None]
    d = d / np.linalg.norm(d, axis=-1, keepdims=True)
    rays = np.zeros(w * h, _ffi.RAY)
    rays["origin"] = np.array(u.position, np.float32)
    rays["dir"] = d.reshape(-1, 3).astype(np.float32)
    return rays


def random_rays(bmin, bmax, n: int, seed: int) -> np.ndarray:
    """Rays from random points of the (slightly grown) bounds in random directions,
    a share of them axis-aligned (zero direction components)."""
    rng = np.random.default_rng(seed)
    bmin = np.asarray(bmin, np.float64)
    bmax = np.asarray(bmax, np.float64)
    ext = bmax - bmin
    o = bmin - 0.1 * ext + rng.random((n, 3)) * 1.2 * ext
    d = rng.normal(size=(n, 3))
    k = n // 8
    d[:k] = 0.0
    d[np.arange(k), rng.integers(0, 3, k)] = rng.choice([-1.0, 1.0], k)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(n, _ffi.RAY)
    rays["origin"] = o.astype(np.float32)
    rays["dir"] = d.astype(np.float32)
    return rays


def face_rays(nodes: np.ndarray, n: int, seed: int) -> np.ndarray:
    """Rays that lie in the plane of a BVH box face: the origin has one
    coordinate exactly on a node's min or max, and the direction is exactly zero
    on that axis (sometimes on a second axis too), so 1/rd is infinite there and
    the slab test meets 0 * inf. Flat boxes (axis-aligned quads) make these the
    rays whose box decisions depend on the NaN rule of min/max."""
    rng = np.random.default_rng(seed)
    k = rng.integers(0, len(nodes), n)
    lo = nodes["bbox_min"][k].astype(np.float32)
    hi = nodes["bbox_max"][k].astype(np.float32)
    a = rng.integers(0, 3, n)
    t = rng.random((n, 3)).astype(np.float32)
    o = lo + (hi - lo) * t
    on_max = rng.random(n) < 0.5
    o[np.arange(n), a] = np.where(on_max, hi[np.arange(n), a], lo[np.arange(n), a])
    d = rng.normal(size=(n, 3))
    d[np.arange(n), a] = 0.0
    b = (a + rng.integers(1, 3, n)) % 3
    two = rng.random(n) < 0.3
    d[np.arange(n)[two], b[two]] = 0.0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(n, _ffi.RAY)
    rays["origin"] = o
    rays["dir"] = d.astype(np.float32)
    assert (rays["dir"][np.arange(n), a] == 0).all()
    return rays


def first_diff(a: np.ndarray, b: np.ndarray) -> str:
    bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
    if len(bad) == 0:
        return "identical"
    i = tuple(bad[0])
    return f"{len(bad)} differing values, first at {i}: {a[i]!r} vs {b[i]!r}"


def icosphere(levels, seed):
    """A closed triangle mesh: an icosahedron subdivided `levels` times, its
    vertices on the unit sphere with a 2% seeded radial jitter (so no two faces
    are coplanar). Shared edges are stored once per triangle with the same
    vertex indices. Returns (float32 vertices, uint32 faces)."""
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = [np.array(x, float) / np.linalg.norm(x) for x in v]
    for _ in range(levels):
        mid = {}
        nf = []
        for a, b, c in f:
            ids = []
            for i, j in ((a, b), (b, c), (c, a)):
                key = (min(i, j), max(i, j))
                if key not in mid:
                    m = v[i] + v[j]
                    v.append(m / np.linalg.norm(m))
                    mid[key] = len(v) - 1
                ids.append(mid[key])
            nf += [(a, ids[0], ids[2]), (b, ids[1], ids[0]), (c, ids[2], ids[1]), (ids[0], ids[1], ids[2])]
        f = nf
    rng = np.random.default_rng(seed)
    V = np.array(v) * (1 + 0.02 * rng.standard_normal((len(v), 1)))
    return V.astype(np.float32), np.array(f, np.uint32)
