"""Pins of the oracle's camera-path and photon-emission branches against closed
forms derived by hand from the reference's WGSL (CPU only).

Each expected value below follows from the reference's estimator, worked out
on a scene where the integral has a closed form (or a 1-D/2-D quadrature of
the published formulas); none re-runs the oracle's own code path:

* furnace plane: a Lambert plane under a uniform sky, seen at normal incidence
  (gather.wgsl:299-537): the diffuse lobe adds the sky twice (the environment
  term at the hit, gather.wgsl:450-452, and the bounce ray's miss), the
  specular lobe survives with the mirror half of the VNDF select
  (gather.wgsl:255) at weight F(1) G1(1) / spec_prob: E = (2 rho + 0.02) S;
* emissive mirror hall: Russian roulette (gather.wgsl:533-537) and the depth
  cut: E = Le * sum_{b<D} 0.25^b;
* sample_spec_dir (gather.wgsl:230-256): the half-vector distribution of the
  visible-normal sampling of Heitz 2018 in its unit-roughness configuration
  (the reference omits the unstretch step) and the mirror fallback taken when
  the sample is above the surface;
* area-light NEE (gather.wgsl:172-203): the self-shadow of the offset shadow
  ray leaves a strip of the light of width d / (1 - 0.01 / h) - d;
* the firefly clamp at 4 spp (gather.wgsl:548-579): only 0, the clamped
  1-of-4 value, E/2, 3E/4 and E can occur;
* photon emission per light kind (photon_trace.wgsl:112-175).
"""
import numpy as np
import pytest
from scipy import stats

import rustracer_amd as rt
from rustracer_amd import _ffi

from . import oracle_ffi as orc
from .helpers import image_params

F = np.float32


def _plane(s, y, half, material, up=True):
    m = s.add_material(material)
    if up:
        p = [[-half, y, half], [half, y, half], [half, y, -half], [-half, y, -half]]
        n = [[0, 1, 0]] * 4
    else:
        p = [[-half, y, -half], [half, y, -half], [half, y, half], [-half, y, half]]
        n = [[0, -1, 0]] * 4
    s.add_mesh(p, [0, 1, 2, 0, 2, 3], m, normals=n)


def _frame_means(scene, cam_pos, look, max_depth, frames, w=16, h=16, spp=16):
    """Mean rgb of each frame (pixels and samples averaged), frames 1..frames."""
    cam = rt.camera(cam_pos, look, up=(0.0, 0.0, -1.0), fov_degrees=0.01)
    cfg = rt.RenderConfig(width=w, height=h, spp=spp, ambient=0.0, enable_denoise=False, max_depth=max_depth)
    u = rt.camera_to_uniform(cam, w, h)
    o = orc.OracleScene.from_scene(scene)
    out = []
    for k in range(1, frames + 1):
        frame, _, _, _ = o.gather(u, image_params(cfg, k), max_depth=max_depth)
        out.append(frame[..., :3].reshape(-1, 3).mean(axis=0, dtype=np.float64))
    return np.array(out)


def _zscore(samples, expected):
    m = samples.mean(axis=0)
    se = samples.std(axis=0, ddof=1) / np.sqrt(len(samples))
    return (m - expected) / np.maximum(se, 1e-12), m


@pytest.mark.parametrize("depth", [1, 3])
def test_furnace_plane_under_uniform_sky(depth):
    rho = np.array([0.5, 0.3, 0.2])
    s = rt.Scene()
    _plane(s, 0.0, 50.0, rt.make_material(0, albedo=tuple(rho), roughness=0.5, metallic=0.0))
    s.add_light(_ffi.LIGHT_ENVIRONMENT, vec_a=(1, 1, 1), vec_b=(1, 1, 1), intensity=1.0)
    means = _frame_means(s, (0.2, 2.0, 0.1), (0.2, 0.0, 0.1), depth, frames=40)
    # depth 1: only the environment term of the diffuse lobe, (1 - q) dscale S = rho S;
    # depth >= 2: the bounce ray's miss adds rho S again, and the specular lobe's
    # mirror half adds q * 0.5 * (F(1) / q) S = 0.02 S (F(1) = F0 = 0.04)
    expected = rho if depth == 1 else 2.0 * rho + 0.02
    z, m = _zscore(means, expected)
    assert np.all(np.abs(z) < 5.0), (m, expected, z)


@pytest.mark.parametrize("depth", [1, 2, 4])
def test_emissive_mirror_hall_russian_roulette_and_depth_cut(depth):
    """Two facing emissive metal planes (albedo 0.5, metallic 1): spec_prob is
    clamped to 0.999, each bounce survives the VNDF select with probability
    1/2 at normal incidence and carries F(1) G1(1) / 0.999 = 0.5 / 0.999, so the
    expected throughput falls by 0.25 per bounce whatever Russian roulette does;
    emission is added at each of the first `depth` hits."""
    s = rt.Scene()
    metal = rt.make_material(1, albedo=(0.5, 0.5, 0.5), roughness=0.3, metallic=1.0, emissive=(1.0, 1.0, 1.0))
    _plane(s, 1.0, 50.0, metal, up=True)
    _plane(s, 2.0, 50.0, metal, up=False)
    means = _frame_means(s, (0.3, 1.5, 0.2), (0.3, 0.0, 0.2), depth, frames=40)
    expected = sum(0.25 ** b for b in range(depth))
    z, m = _zscore(means, np.full(3, expected))
    assert np.all(np.abs(z) < 5.0), (m, expected, z)
    if depth == 1:
        assert np.all(means == 1.0)  # one hit, no randomness in what it adds


def _basis_z(n):
    up = np.array([0.0, 1.0, 0.0]) if abs(n[1]) < 0.999 else np.array([1.0, 0.0, 0.0])
    t = np.cross(up, n)
    t /= np.linalg.norm(t)
    return t, np.cross(n, t)


@pytest.mark.parametrize("theta_deg,a", [(60.0, 0.25), (35.0, 0.6), (75.0, 0.1)])
def test_sample_spec_dir_distribution(theta_deg, a):
    n = np.array([0.2, 1.0, 0.1])
    n /= np.linalg.norm(n)
    t, b = _basis_z(n)
    th = np.radians(theta_deg)
    v = np.cos(th) * n + np.sin(th) * (0.8 * t + 0.6 * b)  # outgoing (view) direction
    rd = (-v).astype(F)
    nf = n.astype(F)
    N = 200_000
    out = orc.sample_spec_dirs(rd, nf, a, N, seed=1234 + int(theta_deg)).astype(np.float64)
    rd64, n64 = rd.astype(np.float64), nf.astype(np.float64)
    mirror = rd64 - 2.0 * (rd64 @ n64) * n64
    is_mirror = np.linalg.norm(out - mirror, axis=1) < 1e-5
    # analytic: in the local frame (t, b, n) the half vector h has density
    # max(0, h . vh) / (pi (1 + vh_z) / 2) on the upper hemisphere, with
    # vh = normalize(a v_x, a v_y, max(v_z, 1e-4)); out = reflect(rd, h),
    # replaced by the mirror direction when out . n > 0
    vl = np.array([t @ -rd64, b @ -rd64, n64 @ -rd64])
    vh = np.array([a * vl[0], a * vl[1], max(vl[2], 1e-4)])
    vh /= np.linalg.norm(vh)
    nc, nphi = 1200, 2400
    cz = (np.arange(nc) + 0.5) / nc
    phi = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    CZ, PH = np.meshgrid(cz, phi, indexing="ij")
    S = np.sqrt(1 - CZ ** 2)
    H = np.stack([S * np.cos(PH), S * np.sin(PH), CZ], -1)
    pdf = np.maximum(H @ vh, 0.0) / (np.pi * (1 + vh[2]) / 2)
    dw = (1.0 / nc) * (2 * np.pi / nphi)
    assert abs(pdf.sum() * dw - 1.0) < 1e-3
    hw = H[..., 0:1] * t + H[..., 1:2] * b + H[..., 2:3] * n64
    o = rd64 - 2.0 * (hw @ rd64)[..., None] * hw
    above = (o @ n64) > 0
    p_mirror = (pdf * above).sum() * dw
    se = np.sqrt(p_mirror * (1 - p_mirror) / N)
    assert abs(is_mirror.mean() - p_mirror) < 5 * se, (is_mirror.mean(), p_mirror)
    assert 0.05 < p_mirror < 0.95 and abs(p_mirror - 0.5) > 10 * se  # the select's side is visible
    # histogram of the recovered half vectors of the non-mirror samples
    # (out - rd is parallel to h; h and -h reflect alike, and the sample lies
    # in the upper hemisphere)
    hs = out[~is_mirror] - rd64
    hs /= np.linalg.norm(hs, axis=1, keepdims=True)
    hl = np.stack([hs @ t, hs @ b, hs @ n64], -1)
    hl *= np.where(hl[:, 2:3] < 0, -1.0, 1.0)
    kc, kp = 10, 16
    ci = np.minimum((hl[:, 2] * kc).astype(int), kc - 1)
    pi_ = np.minimum(((np.arctan2(hl[:, 1], hl[:, 0]) % (2 * np.pi)) / (2 * np.pi) * kp).astype(int), kp - 1)
    obs = np.bincount(ci * kp + pi_, minlength=kc * kp).astype(float)
    w = pdf * (~above) * dw
    exp = w.reshape(kc, nc // kc, kp, nphi // kp).sum(axis=(1, 3)).ravel() * N
    keep = exp > 20
    chi2 = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    assert obs[~keep].sum() <= exp[~keep].sum() + 5 * np.sqrt(exp[~keep].sum() + 1) + 5
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-5, (chi2, keep.sum())


def _spec_ggx(wo, n, wi, a, f0):
    """The GGX lobe of gather.wgsl:214-227 (Walter et al. 2007 D, Smith-Schlick G, Schlick F)."""
    ndl, ndv = wi @ n, wo @ n
    if ndl <= 0 or ndv <= 0:
        return 0.0
    h = (wo + wi) / np.linalg.norm(wo + wi)
    ndh = max(h @ n, 0.0)
    a2 = a * a
    dd = ndh * ndh * (a2 - 1) + 1
    D = a2 / (np.pi * dd * dd)
    c = max(wo @ h, 0.0)
    Fr = f0 + (1 - f0) * (1 - c) ** 5
    g1 = lambda x: x / (x * (1 - a / 2) + a / 2)
    return Fr * D * g1(ndl) * g1(ndv) / (4 * ndl * ndv)


def test_area_light_nee_self_shadow_strip():
    """A down-facing light quad at height h = 0.6, x in [x0, x1], over a Lambert
    plane; the camera sees the plane at the origin. The shadow ray starts at
    hp + 0.01 n and stops at r - 0.001, so it reaches the light's plane at
    lp shrunk towards the origin by (1 - 0.01/h), inside the light unless
    x < x0 / (1 - 0.01/h): NEE (depth 1) integrates
    (rho/pi + spec_ggx) Le cos^2 / (r^2 + 0.01) over that strip only."""
    h, x0, x1, zh, Le, rho = 0.6, 0.7, 1.3, 0.3, 10.0, 0.6
    s = rt.Scene()
    _plane(s, 0.0, 20.0, rt.make_material(0, albedo=(rho, rho, rho), roughness=1.0, metallic=0.0))
    light = s.add_material(rt.make_material(3, albedo=(0, 0, 0), emissive=(Le, Le, Le)))
    first = s.counts()["triangles"]
    s.add_mesh([[x0, h, -zh], [x1, h, -zh], [x1, h, zh], [x0, h, zh]], [0, 1, 2, 0, 2, 3], light,
               normals=[[0, -1, 0]] * 4)
    for k in range(2):
        s.add_light(_ffi.LIGHT_AREA, color=(Le, Le, Le), intensity=1.0, triangle_index=first + k)
    tris = s.flatten_triangles()
    for k in range(2):  # the light faces the plane
        v0, v1, v2 = (np.asarray(tris[first + k][c], float) for c in ("v0", "v1", "v2"))
        assert np.cross(v1 - v0, v2 - v0)[1] < 0
    means = _frame_means(s, (0.0, 1.5, 0.0), (0.0, 0.0, 0.0), 1, frames=160)
    xs = x0 / (1.0 - 0.01 / h)
    gx, gw = np.polynomial.legendre.leggauss(48)
    X = x0 + (xs - x0) * (gx + 1) / 2
    Z = zh * gx
    wo, nrm = np.array([0.0, 1.0, 0.0]), np.array([0.0, 1.0, 0.0])
    total = 0.0
    full = 0.0
    for i, x in enumerate(X):
        for j, z in enumerate(Z):
            r2 = x * x + h * h + z * z
            ld = np.array([x, h, z]) / np.sqrt(r2)
            g = (rho / np.pi + _spec_ggx(wo, nrm, ld, 1.0, 0.04)) * Le * ld[1] ** 2 / (r2 + 0.01)
            total += gw[i] * gw[j] * g
    total *= (xs - x0) / 2 * zh
    z, m = _zscore(means, np.full(3, total))
    assert np.all(np.abs(z) < 5.0), (m, total, z)
    # without the self-shadow the same integral over the whole light is ~60x larger
    for i, x in enumerate(x0 + (x1 - x0) * (gx + 1) / 2):
        for j, zz in enumerate(Z):
            r2 = x * x + h * h + zz * zz
            ld = np.array([x, h, zz]) / np.sqrt(r2)
            full += gw[i] * gw[j] * (rho / np.pi + _spec_ggx(wo, nrm, ld, 1.0, 0.04)) * Le * ld[1] ** 2 / (r2 + 0.01)
    full *= (x1 - x0) / 2 * zh
    assert full > 20 * total


def test_firefly_clamp_values_at_four_samples():
    """Pixels on the edge of an emitter against the black background see k of 4
    samples on it. The clamp of gather.wgsl:548-579 keeps k = 2, 3, 4 (E k/4)
    and cuts k = 1 to E * thresh / (4 L), thresh = 3 * 1e-5 (median 0, zero
    spread); E/4 never occurs."""
    E = np.array([2.0, 1.5, 1.0], F)
    s = rt.Scene()
    m = s.add_material(rt.make_material(3, albedo=(0, 0, 0), emissive=tuple(E)))
    s.add_mesh([[0, -5, -3], [5, -5, -3], [5, 5, -3], [0, 5, -3]], [0, 1, 2, 0, 2, 3], m, normals=[[0, 0, 1]] * 4)
    w, hgt = 8, 4
    # sample positions of pixel x span [x - 0.5, x + 0.5) (gather.wgsl:292-294),
    # so the image centre (the emitter's edge, x = 0) runs through column w/2
    cam = rt.camera((0.0, 0.0, 0.0), (0.0, 0.0, -1.0), fov_degrees=40)
    cfg = rt.RenderConfig(width=w, height=hgt, spp=4, ambient=0.0, enable_denoise=False, max_depth=1)
    u = rt.camera_to_uniform(cam, w, hgt)
    o = orc.OracleScene.from_scene(s)
    vals = []
    for k in range(1, 121):
        frame, mom, _, _ = o.gather(u, image_params(cfg, k), max_depth=1)
        vals.append(frame[:, w // 2, :3])
    vals = np.concatenate(vals).astype(np.float64)
    L = float((F(E[0]) * F(0.2126) + F(E[1]) * F(0.7152)) + F(E[2]) * F(0.0722))
    thresh = float(F(F(0.25 * 0.0) + F(1e-5)) * F(3.0))
    allowed = [np.zeros(3), E * thresh / L / 4, E / 2, 3 * E / 4, E.astype(np.float64)]
    kind = np.full(len(vals), -1)
    for i, a in enumerate(allowed):
        close = np.all(np.abs(vals - a) <= 1e-6 * np.maximum(np.abs(a), 1e-3), axis=1)
        kind[close] = i
    assert (kind >= 0).all(), vals[kind < 0][:5]
    counts = np.bincount(kind, minlength=5)
    assert counts[1] > 0 and counts[2] > 0 and counts[3] > 0, counts  # 1-of-4 pixels exist and are clamped
    assert not np.any(np.all(np.abs(vals - E / 4) < 1e-4, axis=1))


def _emission_scene():
    s = rt.Scene()
    m = s.add_material(rt.make_material(0))
    first = s.counts()["triangles"]
    s.add_mesh([[-1.0, 3.0, -0.5], [1.5, 3.2, -0.3], [0.2, 2.8, 1.0]], [0, 1, 2], m)
    s.add_mesh([[-6, -1, -6], [6, -1, -6], [6, 5, 6]], [0, 1, 2], m)  # bounds for the grid
    s.add_light(_ffi.LIGHT_POINT, vec_a=(0.3, 1.0, -0.2), color=(1.0, 0.9, 0.8), intensity=5.0)
    s.add_light(_ffi.LIGHT_DIRECTIONAL, vec_a=(0.3, -1.0, 0.2), color=(0.7, 0.8, 0.9), intensity=2.0)
    s.add_light(_ffi.LIGHT_AREA, color=(3.0, 2.0, 1.0), intensity=1.5, triangle_index=first)
    s.add_light(_ffi.LIGHT_ENVIRONMENT, vec_a=(0.2, 0.4, 0.9), vec_b=(0.9, 0.8, 0.6), intensity=0.7)
    return s, first


def test_photon_emission_per_light_kind():
    s, tri_index = _emission_scene()
    o = orc.OracleScene.from_scene(s)
    lo, hi = s.bounds()
    cfg = rt.RenderConfig(enable_photon_map=True, photon_count=1000)
    gp, _ = rt.grid_params_for(cfg, lo, hi)
    lights = s.lights_gpu()
    N = 60_000
    rec = {k: [] for k in range(4)}
    for i in range(N):
        r = o.photon_emit(gp, i)
        assert r is not None
        origin, d, power, li = r
        rec[int(lights[li]["kind"])].append((origin.astype(float), d.astype(float), power.astype(float)))
    counts = np.array([len(rec[k]) for k in range(4)])
    assert stats.chisquare(counts).pvalue > 1e-4, counts  # the light pick is uniform
    ks = lambda x, cdf: stats.kstest(x, cdf).pvalue

    # point light: origin at the light; the reference normalises
    # (2u r, 2v r, 1 - 2s), r = sqrt(1 - (1 - 2s)^2), s = u^2 + v^2 uniform on
    # [0, 1): dir.z = f(s) = (1 - 2s) / sqrt(16 s^2 (1 - s) + (1 - 2s)^2), which
    # decreases from 1 to -1, so P(dir.z <= z) = 1 - f^-1(z); azimuth uniform
    O = np.array([r[0] for r in rec[0]])
    D = np.array([r[1] for r in rec[0]])
    assert np.all(O == np.array([0.3, 1.0, -0.2], F))
    sg = np.linspace(0, 1, 200_001)
    fz = (1 - 2 * sg) / np.sqrt(16 * sg ** 2 * (1 - sg) + (1 - 2 * sg) ** 2)
    assert np.all(np.diff(fz) < 0)
    cdf_z = lambda z: 1.0 - np.interp(z, fz[::-1], sg[::-1])
    assert ks(D[:, 2], cdf_z) > 1e-4
    assert ks(D[:, 2], stats.uniform(-1, 2).cdf) < 1e-6  # it is not the uniform sphere
    assert ks(np.arctan2(D[:, 1], D[:, 0]) % (2 * np.pi), stats.uniform(0, 2 * np.pi).cdf) > 1e-4
    assert np.all(np.array([r[2] for r in rec[0]]) == (np.array([1.0, 0.9, 0.8], F) * F(5.0)))

    # directional: origin = u t + v b - 100 d with u, v uniform on [-5, 5), dir = d
    d = np.array([0.3, -1.0, 0.2])
    up = np.array([0.0, 1.0, 0.0]) if abs(d[1]) < 0.999 else np.array([1.0, 0.0, 0.0])
    tt = np.cross(up, d)
    tt /= np.linalg.norm(tt)
    bb = np.cross(d, tt)
    O = np.array([r[0] for r in rec[1]]) + 100 * d
    assert np.all(np.array([r[1] for r in rec[1]]) == d.astype(F))
    assert np.allclose(O @ d, 0, atol=1e-3)
    for axis in (tt, bb):
        assert ks(O @ axis / (axis @ axis), stats.uniform(-5, 10).cdf) > 1e-4

    # area light: origin uniform on the triangle (each barycentric is Beta(1, 2)),
    # direction cosine-distributed about the triangle's normal
    tri = s.flatten_triangles()[tri_index]
    v0, v1, v2 = (np.asarray(tri[c], float) for c in ("v0", "v1", "v2"))
    nrm = np.cross(v1 - v0, v2 - v0)
    nrm /= np.linalg.norm(nrm)
    O = np.array([r[0] for r in rec[2]])
    D = np.array([r[1] for r in rec[2]])
    M = np.stack([v0, v1, v2], 1)
    lam = np.linalg.lstsq(np.vstack([M, np.ones(3)]), np.vstack([O.T, np.ones(len(O))]), rcond=None)[0].T
    assert np.all(lam > -1e-4) and np.allclose(lam.sum(1), 1.0)
    for k in range(3):
        assert ks(lam[:, k], lambda x: 1 - (1 - np.clip(x, 0, 1)) ** 2) > 1e-4
    c = D @ nrm
    assert np.all(c > -1e-6)
    assert ks(c ** 2, stats.uniform(0, 1).cdf) > 1e-4
    assert np.all(np.array([r[2] for r in rec[2]]) == (np.array([3.0, 2.0, 1.0], F) * F(1.5)))

    # environment: from the dome of radius R around the grid's centre, towards
    # it; cosine-distributed about +y; power = mix(zenith, horizon, sqrt(t)) I,
    # t = dir.y / 2 + 1/2 (zenith weight at the bottom: photon_trace.wgsl:167-168)
    span = np.array([gp.grid_x, gp.grid_y, gp.grid_z], float) * gp.cell_size
    centre = np.array([gp.room_min_x, gp.room_min_y, gp.room_min_z]) + span / 2
    R = span.max() * 1.2 + 0.05
    O = np.array([r[0] for r in rec[3]])
    D = np.array([r[1] for r in rec[3]])
    P = np.array([r[2] for r in rec[3]])
    e = -D
    assert np.allclose(np.linalg.norm(O - centre, axis=1), R, rtol=1e-4)
    assert np.allclose((O - centre) / R, e, atol=1e-4)
    assert np.all(e[:, 1] >= -1e-6)
    assert ks(e[:, 1] ** 2, stats.uniform(0, 1).cdf) > 1e-4
    zen, hor = np.array([0.2, 0.4, 0.9]), np.array([0.9, 0.8, 0.6])
    tw = np.sqrt(np.clip(e[:, 1] * 0.5 + 0.5, 0, 1))[:, None]
    assert np.allclose(P, (zen * (1 - tw) + hor * tw) * 0.7, rtol=2e-5, atol=1e-6)


@pytest.mark.parametrize("depth", [1, 2])
def test_alpha_cutout_pass_through(depth):
    """An emissive cutout quad (2x2 alpha checker, bilinear/repeat) in front of
    a second emitter: a camera ray meeting an opaque texel sees the front
    emission E2; at a transparent texel (alpha < 0.5) the path continues 0.002
    behind it and the bounce count advances (gather.wgsl:315-323), so it sees
    the back emitter E1 with depth 2 and nothing with depth 1."""
    E1, E2 = np.array([1.0, 2.0, 3.0], F), np.array([4.0, 0.5, 0.25], F)
    s = rt.Scene()
    back = s.add_material(rt.make_material(3, albedo=(0, 0, 0), emissive=tuple(E1)))
    s.add_mesh([[-4, -4, -5], [4, -4, -5], [4, 4, -5], [-4, 4, -5]], [0, 1, 2, 0, 2, 3], back,
               normals=[[0, 0, 1]] * 4)
    tex = np.ones((2, 2, 4), np.float32)
    tex[0, 1, 3] = tex[1, 0, 3] = 0.0  # alpha checker: texel (x, y) opaque when x == y
    layer = s.add_texture(0, tex)
    front = rt.make_material(3, albedo=(0, 0, 0), emissive=tuple(E2))
    front["tex_albedo"] = layer
    front["flags"] = 1  # RT_MAT_FLAG_ALPHA_CUTOUT
    m = s.add_material(front)
    s.add_mesh([[-1, -1, -3], [1, -1, -3], [1, 1, -3], [-1, 1, -3]], [0, 1, 2, 0, 2, 3], m,
               normals=[[0, 0, 1]] * 4, uvs=[[0, 0], [1, 0], [1, 1], [0, 1]])
    w = h = 64
    cam = rt.camera((0.0, 0.0, 0.0), (0.0, 0.0, -1.0), fov_degrees=40)
    cfg = rt.RenderConfig(width=w, height=h, spp=4, ambient=0.0, enable_denoise=False, max_depth=depth)
    frame, _, _, _ = orc.OracleScene.from_scene(s).gather(rt.camera_to_uniform(cam, w, h), image_params(cfg, 1),
                                                          max_depth=depth)
    # pixel footprint -> quad uv (u = (x + 1) / 2, v = (y + 1) / 2 at distance 3)
    fy = np.tan(np.radians(20))
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    opaque = np.ones((h, w), bool)
    clear = np.ones((h, w), bool)
    for jx in (-0.5, 0.0, 0.5):
        for jy in (-0.5, 0.0, 0.5):
            u = (3 * ((xs + jx) * 2 / w - 1) * fy + 1) / 2
            v = (3 * (1 - (ys + jy) * 2 / h) * fy + 1) / 2
            inside = (u > 0.02) & (u < 0.98) & (v > 0.02) & (v < 0.98)
            # near a texel centre (0.25 / 0.75): the bilinear alpha is that texel's
            near = (np.abs((u % 0.5) - 0.25) < 0.1) & (np.abs((v % 0.5) - 0.25) < 0.1)
            tx, ty = (u > 0.5).astype(int), (v > 0.5).astype(int)
            opaque &= inside & near & (tx == ty)
            clear &= inside & near & (tx != ty)
    assert opaque.sum() > 50 and clear.sum() > 50
    assert np.all(frame[opaque][:, :3] == E2)
    want = E1 if depth == 2 else np.zeros(3, F)
    assert np.all(frame[clear][:, :3] == want)


def _tangent_views(positions, normal, uvs, cam_pos, look, up):
    """Debug views 19 and 20 (tangent * 0.5 + 0.5, bitangent * 0.5 + 0.5) at
    the centre pixels of a quad filling a 32x32 image. At 9 spp (above the
    firefly clamp's 8, which would scale a debug value by a luminance sum that
    debug values do not enter) the frame holds the last sample's value / 9."""
    s = rt.Scene()
    m = s.add_material(rt.make_material(0))
    s.add_mesh(positions, [0, 1, 2, 0, 2, 3], m, normals=[normal] * 4, uvs=uvs)
    cam = rt.camera(cam_pos, look, up=up, fov_degrees=30)
    out = []
    for view in (19, 20):
        cfg = rt.RenderConfig(width=32, height=32, spp=9, enable_denoise=False, max_depth=1, debug_view=view)
        frame, _, _, _ = orc.OracleScene.from_scene(s).gather(rt.camera_to_uniform(cam, 32, 32),
                                                              image_params(cfg, 1), max_depth=1)
        c = frame[12:20, 12:20, :3].reshape(-1, 3).astype(np.float64) * 9.0
        assert np.ptp(c, axis=0).max() < 1e-5  # a flat quad: one frame everywhere
        out.append(c[0] * 2.0 - 1.0)
    return out


QUAD_Z = [[-1, -1, -3], [1, -1, -3], [1, 1, -3], [-1, 1, -3]]  # normal +z, towards a camera at the origin


@pytest.mark.parametrize("uvs,t_exp,b_exp", [
    # u along +x, v along +y: T = +x, B = n x T = +y
    ([[0, 0], [1, 0], [1, 1], [0, 1]], (1, 0, 0), (0, 1, 0)),
    # u mirrored (along -x): T = -x; n x T = -y points against dP/dv, so B flips to +y
    ([[1, 0], [0, 0], [0, 1], [1, 1]], (-1, 0, 0), (0, 1, 0)),
    # u along +y, v along -x: T = +y, B = n x T = -x (with dP/dv)
    ([[0, 1], [0, 0], [1, 0], [1, 1]], (0, 1, 0), (-1, 0, 0)),
    # sheared map u = x, v = x + y: dP/du = (1, -1, 0) / |.|, B = n x T
    ([[-1, -2], [1, 0], [1, 2], [-1, 0]], (0.5 ** 0.5, -(0.5 ** 0.5), 0), (0.5 ** 0.5, 0.5 ** 0.5, 0)),
    # degenerate uvs (det = 0): T falls back to n x (1, 0, 0) = +y, B = -(n x T) = +x (dot with b = 0 is not > 0)
    ([[0, 0], [0, 0], [0, 0], [0, 0]], (0, 1, 0), (1, 0, 0)),
])
def test_tangent_frame_closed_forms(uvs, t_exp, b_exp):
    """The normal-map tangent frame (common.wgsl:211-230, used at gather.wgsl:
    335-340): the tangent is the surface direction of increasing u projected
    off the normal, the bitangent n x T turned to point along increasing v;
    with no uv parameterisation the first fallback axis is used. Closed forms
    for a quad facing the camera under five uv layouts."""
    t, b = _tangent_views(QUAD_Z, (0, 0, 1), uvs, (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0))
    assert np.allclose(t, t_exp, atol=1e-5), t
    assert np.allclose(b, b_exp, atol=1e-5), b


def test_tangent_frame_second_fallback_axis():
    """Degenerate uvs on a quad whose normal is +x: n x (1, 0, 0) vanishes, so
    the tangent is n x (0, 1, 0) = +z and the bitangent -(n x T) = +y."""
    quad_x = [[-3, -1, 1], [-3, -1, -1], [-3, 1, -1], [-3, 1, 1]]  # normal +x
    t, b = _tangent_views(quad_x, (1, 0, 0), [[0, 0]] * 4, (0.0, 0.0, 0.0), (-1.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    assert np.allclose(t, (0, 0, 1), atol=1e-5), t
    assert np.allclose(b, (0, 1, 0), atol=1e-5), b


def test_tangent_frame_projects_off_a_tilted_shading_normal():
    """Vertex normals tilted to n = (0.6, 0, 0.8) on the quad facing +z: the
    tangent is dP/du = +x with its component along n removed,
    (1, 0, 0) - 0.6 n = (0.64, 0, -0.48), normalised (0.8, 0, -0.6); the
    bitangent n x T = +y points along dP/dv."""
    t, b = _tangent_views(QUAD_Z, (0.6, 0.0, 0.8), [[0, 0], [1, 0], [1, 1], [0, 1]], (0.0, 0.0, 0.0),
                          (0.0, 0.0, -1.0), (0.0, 1.0, 0.0))
    assert np.allclose(t, (0.8, 0.0, -0.6), atol=1e-5), t
    assert np.allclose(b, (0.0, 1.0, 0.0), atol=1e-5), b
