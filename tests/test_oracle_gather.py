"""Oracle camera paths (gather.wgsl:258-594) against closed forms. CPU only.

* emission only: every pixel that sees an emitter returns its radiance exactly;
* camera + G-buffer: first-hit depth and normal of a plane;
* next-event estimation toward a point light and a directional light over a
  Lambert plane: the per-pixel mean converges to
  rho*E/pi + spec_ggx(wo, n, wi, a=1, F0=0.04)*E  (the lobe choice is unbiased),
  integrated over the pixel footprint; checked with per-pixel z-scores.
"""
import numpy as np
import pytest

import rustracer_amd as rt
from rustracer_amd import _ffi

from . import oracle_ffi as orc
from .helpers import image_params

F = np.float32


def _quad_scene(y_plane=None, z_plane=None, material=None, half=4.0):
    s = rt.Scene()
    m = s.add_material(material)
    if z_plane is not None:
        p = [[-half, -half, z_plane], [half, -half, z_plane], [half, half, z_plane], [-half, half, z_plane]]
        n = [[0, 0, 1]] * 4
    else:
        p = [[-half, y_plane, half], [half, y_plane, half], [half, y_plane, -half], [-half, y_plane, -half]]
        n = [[0, 1, 0]] * 4
    s.add_mesh(p, [0, 1, 2, 0, 2, 3], m, normals=n)
    return s


def _cfg(w, h, spp, ambient=0.0):
    return rt.RenderConfig(width=w, height=h, spp=spp, ambient=ambient, enable_denoise=False)


def test_emission_only_exact():
    s = _quad_scene(z_plane=-3.0, material=rt.make_material(3, albedo=(1, 1, 1), emissive=(2.5, 2.5, 2.5)),
                    half=1.5)
    cam = rt.camera([0, 0, 0], [0, 0, -1], fov_degrees=60)
    w, h = 40, 30
    cfg = _cfg(w, h, 4)
    u = rt.camera_to_uniform(cam, w, h)
    o = orc.OracleScene.from_scene(s)
    frame, mom, gb, _ = o.gather(u, image_params(cfg, 1))
    # pixel footprints entirely inside / outside the quad (|x|, |y| <= 1.5 at distance 3)
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    fy = np.tan(np.radians(30))
    fx = fy * w / h
    inside = np.ones((h, w), bool)
    outside = np.ones((h, w), bool)
    for jx in (-0.5, 0.5):
        for jy in (-0.5, 0.5):
            x = 3 * ((xs + jx) * 2 / w - 1) * fx
            y = 3 * (1 - (ys + jy) * 2 / h) * fy
            cin = (np.abs(x) < 1.5 - 1e-4) & (np.abs(y) < 1.5 - 1e-4)
            inside &= cin
            outside &= (np.abs(x) > 1.5 + 1e-4) | (np.abs(y) > 1.5 + 1e-4)
    assert inside.sum() > 100 and outside.sum() > 100
    assert (frame[inside][:, :3] == F(2.5)).all()
    assert (frame[outside][:, :3] == 0).all() and (gb[outside][:, 3] == F(1e30)).all()
    edge = ~inside & ~outside
    assert ((frame[edge][:, :3] >= 0) & (frame[edge][:, :3] <= F(2.5))).all()
    L = (F(2.5) * F(0.2126) + F(2.5) * F(0.7152)) + F(2.5) * F(0.0722)
    assert (mom[inside][:, 0] == L).all() and (mom[inside][:, 1] == L * L).all()


def test_gbuffer_depth_and_normal_of_a_plane():
    s = _quad_scene(z_plane=-3.0, material=rt.make_material(0), half=50)
    cam = rt.camera([0, 0, 0], [0, 0, -1], fov_degrees=70)
    w, h = 32, 24
    u = rt.camera_to_uniform(cam, w, h)
    frame, mom, gb, _ = orc.OracleScene.from_scene(s).gather(u, image_params(_cfg(w, h, 3), 1))
    assert (gb[..., :3] == np.array([0, 0, 1], F)).all()
    # depth = 3 / cos(angle of the ray); bound it over the pixel footprint
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    fx, fy = np.tan(np.radians(35)) * w / h, np.tan(np.radians(35))
    lo = np.full((h, w), np.inf)
    hi = np.zeros((h, w))
    for jx in np.linspace(-0.5, 0.5, 21):
        for jy in np.linspace(-0.5, 0.5, 21):
            ux = (xs + jx) * 2 / w - 1
            uy = 1 - (ys + jy) * 2 / h
            t = 3.0 * np.sqrt(1 + (ux * fx) ** 2 + (uy * fy) ** 2)
            lo, hi = np.minimum(lo, t), np.maximum(hi, t)
    d = gb[..., 3].astype(np.float64)
    assert ((d >= lo * (1 - 1e-5)) & (d <= hi * (1 + 1e-5))).all()


def _spec_ggx(wo, n, wi, a, f0):
    ndl = (n * wi).sum(-1)
    ndv = (n * wo).sum(-1)
    hh = wo + wi
    hh /= np.linalg.norm(hh, axis=-1, keepdims=True)
    ndh = np.maximum((n * hh).sum(-1), 0)
    a2 = a * a
    dd = ndh * ndh * (a2 - 1) + 1
    D = a2 / (np.pi * dd * dd)
    Fr = f0 + (1 - f0) * np.clip(1 - np.maximum((wo * hh).sum(-1), 0), 0, 1) ** 5
    k = a * 0.5
    g1 = lambda x: x / (x * (1 - k) + k)
    out = Fr * D * g1(ndl) * g1(ndv) / (4 * ndl * ndv)
    return np.where((ndl > 0) & (ndv > 0), out, 0.0)


@pytest.mark.parametrize("kind", ["point", "directional"])
def test_nee_converges_to_closed_form(kind):
    rho = 0.5
    s = _quad_scene(y_plane=0.0, material=rt.make_material(0, albedo=(rho, rho, rho), roughness=1.0), half=20)
    if kind == "point":
        s.add_light(_ffi.LIGHT_POINT, vec_a=(0.3, 1.0, -0.2), color=(1, 1, 1), intensity=2.0)
    else:
        s.add_light(_ffi.LIGHT_DIRECTIONAL, vec_a=(0.3, 0.8, 0.2), color=(1, 1, 1), intensity=1.5)
    cam = rt.camera([0, 2.5, 2.0], [0, 0, 0], fov_degrees=50)
    w, h, spp, frames = 16, 12, 16, 24
    cfg = _cfg(w, h, spp)
    u = rt.camera_to_uniform(cam, w, h)
    o = orc.OracleScene.from_scene(s)
    acc = []
    for k in range(1, frames + 1):
        f, _, _, _ = o.gather(u, image_params(cfg, k))
        acc.append(f[..., 0].astype(np.float64))
    acc = np.array(acc)
    mean, se = acc.mean(0), acc.std(0, ddof=1) / np.sqrt(frames)
    # closed form averaged over the pixel footprint (16 x 16 stratified jitters)
    pos, fwd, right, up = (np.array(v, np.float64) for v in (u.position, u.forward, u.right, u.up))
    g = (np.arange(16) + 0.5) / 16 - 0.5
    jx, jy = np.meshgrid(g, g)
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    ux = (xs[..., None, None] + jx) * (2.0 / w) - 1
    uy = 1 - (ys[..., None, None] + jy) * (2.0 / h)
    d = fwd + right * ux[..., None] + up * uy[..., None]
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    t = -pos[1] / d[..., 1]
    hp = pos + d * t[..., None]
    n = np.array([0.0, 1.0, 0.0])
    if kind == "point":
        to = np.array([0.3, 1.0, -0.2]) - hp
        r2 = (to * to).sum(-1)
        wi = to / np.sqrt(r2)[..., None]
        E = 2.0 * np.maximum(wi[..., 1], 0) / (r2 + 0.01)
    else:
        wi = np.broadcast_to(np.array([0.3, 0.8, 0.2]) / np.linalg.norm([0.3, 0.8, 0.2]), hp.shape)
        E = 1.5 * wi[..., 1] * np.ones(hp.shape[:-1])
    val = rho * E / np.pi + _spec_ggx(-d, n, wi, 1.0, 0.04) * E
    expect = val.mean(axis=(-1, -2))
    z = np.abs(mean - expect) / np.maximum(se, 1e-6)
    assert np.median(z) < 1.5 and z.max() < 5.5, (np.median(z), z.max())
    assert abs(mean.sum() / expect.sum() - 1) < 0.01


def test_dielectric_and_sky_on_bounce_misses():
    """A lone glass pane under a uniform sky (zenith = horizon = c, intensity I),
    16 spp (no firefly clamp). gather.wgsl:298-306,386-399: a primary miss is
    black; at the pane a sample reflects with Schlick probability F(cos) and
    then misses, adding c*I, or refracts (throughput * albedo) and misses,
    adding albedo*c*I; nothing else lights the pane. So every pixel inside the
    pane is (k*cI + (16-k)*albedo*cI)/16 for an integer k, the same k in every
    channel, and the k are Binomial(16, F) with F from the pixel's angle."""
    albedo = np.array([0.8, 0.6, 0.4])
    glass = rt.make_material(2, albedo=tuple(albedo), roughness=0.0, metallic=0.0, ior=1.5)
    s = _quad_scene(z_plane=-3.0, material=glass, half=1.5)
    c, intensity = np.array([0.5, 0.7, 0.9]), 2.0
    s.add_light(_ffi.LIGHT_ENVIRONMENT, vec_a=tuple(c), vec_b=tuple(c), intensity=intensity)
    cam = rt.camera([0, 0, 0], [0, 0, -1], fov_degrees=60)
    w, h, spp = 40, 30, 16
    cfg = _cfg(w, h, spp)
    frame, _, _, _ = orc.OracleScene.from_scene(s).gather(rt.camera_to_uniform(cam, w, h), image_params(cfg, 1),
                                                          max_depth=4)
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    fy = np.tan(np.radians(30))
    fx = fy * w / h
    inside = np.ones((h, w), bool)
    outside = np.ones((h, w), bool)
    for jx in (-0.5, 0.5):
        for jy in (-0.5, 0.5):
            x = 3 * ((xs + jx) * 2 / w - 1) * fx
            y = 3 * (1 - (ys + jy) * 2 / h) * fy
            inside &= (np.abs(x) < 1.5 - 1e-4) & (np.abs(y) < 1.5 - 1e-4)
            outside &= (np.abs(x) > 1.5 + 1e-4) | (np.abs(y) > 1.5 + 1e-4)
    assert inside.sum() > 100 and outside.sum() > 100
    assert (frame[outside][:, :3] == 0).all(), "primary misses see black"
    sky = c * intensity
    f = frame[inside][:, :3].astype(np.float64)
    k = (spp * f / sky - spp * albedo) / (1.0 - albedo)  # reflected samples, per channel
    kr = np.round(k[:, 0])
    assert np.abs(k - kr[:, None]).max() < 1e-3, "one integer k explains all three channels"
    assert ((kr >= 0) & (kr <= spp)).all()
    # expected reflections from Schlick's F at each pixel's mean ray (through (px, py))
    x = (xs * 2 / w - 1) * fx
    y = (1 - ys * 2 / h) * fy
    cos = (1.0 / np.sqrt(x * x + y * y + 1.0))[inside]
    f0 = ((1.5 - 1.0) / (1.5 + 1.0)) ** 2
    fr = f0 + (1 - f0) * (1 - cos) ** 5
    mean, var = (spp * fr).sum(), (spp * fr * (1 - fr)).sum()
    assert abs(kr.sum() - mean) < 5 * np.sqrt(var), (kr.sum(), mean, np.sqrt(var))


def test_sky_gradient_seen_through_reflection_and_refraction():
    """gather.wgsl:107-111 sky_color = mix(horizon, zenith, sqrt(clamp(y/2 + 1/2))) * I,
    seen by the reflected ray (y unchanged by the pane at z = const) and the
    refracted ray (y scaled by 1/ior, Snell at a face of normal +z). Row sums of
    the pane's pixels match the expectation with Schlick's F within 5 sigma,
    and swapping zenith and horizon would not."""
    albedo = np.array([0.9, 0.9, 0.9])
    glass = rt.make_material(2, albedo=tuple(albedo), roughness=0.0, metallic=0.0, ior=1.5)
    s = _quad_scene(z_plane=-3.0, material=glass, half=1.5)
    zen, hor, intensity = np.array([0.1, 0.3, 1.0]), np.array([1.0, 0.8, 0.2]), 1.5
    s.add_light(_ffi.LIGHT_ENVIRONMENT, vec_a=tuple(zen), vec_b=tuple(hor), intensity=intensity)
    cam = rt.camera([0, 0, 0], [0, 0, -1], fov_degrees=60)
    w, h, spp = 40, 30, 64
    cfg = _cfg(w, h, spp)
    frame, _, _, _ = orc.OracleScene.from_scene(s).gather(rt.camera_to_uniform(cam, w, h), image_params(cfg, 1),
                                                          max_depth=4)

    def sky(y, z_, h_):
        t = np.clip(y * 0.5 + 0.5, 0.0, 1.0)
        return (h_ * (1 - np.sqrt(t))[..., None] + z_ * np.sqrt(t)[..., None]) * intensity

    fy = np.tan(np.radians(30))
    fx = fy * w / h
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    # the camera ray of pixel (px, py) passes through (px + jx, py + jy), jx, jy in [-1/2, 1/2)
    x = (xs * 2 / w - 1) * fx
    y = (1 - ys * 2 / h) * fy
    inside = (np.abs(3 * x) < 1.4) & (np.abs(3 * y) < 1.4)
    norm = np.sqrt(x * x + y * y + 1.0)
    ry, cos = y / norm, 1.0 / norm
    f0 = 0.04
    fr = f0 + (1 - f0) * (1 - cos) ** 5
    sr, st = sky(ry, zen, hor), albedo * sky(ry / 1.5, zen, hor)
    expect = fr[..., None] * sr + (1 - fr[..., None]) * st
    var = fr[..., None] * (1 - fr[..., None]) * (sr - st) ** 2 / spp + 1e-12
    got = frame[..., :3].astype(np.float64)
    swapped = (fr[..., None] * sky(ry, hor, zen) + (1 - fr[..., None]) * albedo * sky(ry / 1.5, hor, zen))
    rows = [r for r in range(h) if inside[r].sum() > 5]
    assert len(rows) > 10
    for r in rows:
        m = inside[r]
        d = (got[r][m] - expect[r][m]).sum(0)
        sd = np.sqrt(var[r][m].sum(0)) + 2e-4 * m.sum()  # + the sub-pixel spread of y
        assert (np.abs(d) < 5 * sd).all(), (r, d, sd)
    dsw = np.abs((got - swapped)[inside]).mean()
    assert dsw > 20 * np.abs((got - expect)[inside]).mean()
