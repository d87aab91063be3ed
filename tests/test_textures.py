"""Texture arrays and sampling (include/rt_texture.h, used by the kernels and
the oracle): the reference's upload quantisation and layer sizing
(lib.rs:1286-1358) and the bilinear/repeat sampler (lib.rs:508-517), pinned by
closed forms and a float64 brute-force bilinear."""
import numpy as np

import rustracer_amd as rt

from . import oracle_ffi as orc


def test_quantisation_truncates_and_clamps():
    s = rt.Scene()
    vals = np.array([0.0, 0.5, 1.0, 1.5, -1.0, np.nan, 0.999, 0.00392], np.float32)
    img = np.zeros((1, 2, 4), np.float32)
    img.reshape(-1)[:] = vals
    assert s.add_texture(0, img) == 0
    a = s.texture_array(0)
    assert a.shape == (1, 1, 2, 4)
    # (u8)(clamp(c, 0, 1) * 255): 127.5 -> 127, 254.745 -> 254, 0.9996 -> 0, NaN -> 0
    assert a.reshape(-1).tolist() == [0, 127, 255, 255, 0, 0, 254, 0]


def test_layers_take_the_first_texture_size():
    s = rt.Scene()
    s.add_texture(2, np.full((4, 4, 4), 1.0, np.float32))
    s.add_texture(2, np.full((2, 3, 4), 0.5, np.float32))
    s.add_texture(2, np.full((6, 6, 4), 0.2, np.float32))
    a = s.texture_array(2)
    assert a.shape == (3, 4, 4, 4)
    assert (a[0] == 255).all()
    assert (a[1, :2, :3] == 127).all() and (a[1, 2:] == 0).all() and (a[1, :, 3:] == 0).all()
    assert (a[2] == 51).all()  # 0.2*255 = 51.0000008 -> 51 (cropped to 4x4)
    # an empty role is one 1x1 zero layer
    assert s.texture_array(1).shape == (1, 1, 1, 4) and not s.texture_array(1).any()


def _bilinear64(t, u, v, layer):
    layers, h, w, _ = t.shape
    layer = min(max(layer, 0), layers - 1)
    x, y = u * w - 0.5, v * h - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    ax, ay = x - x0, y - y0
    f = t[layer].astype(np.float64) / 255.0
    i0, i1 = int(x0) % w, int(x0 + 1) % w
    j0, j1 = int(y0) % h, int(y0 + 1) % h
    return ((1 - ax) * (1 - ay) * f[j0, i0] + ax * (1 - ay) * f[j0, i1] + (1 - ax) * ay * f[j1, i0]
            + ax * ay * f[j1, i1])


def test_sampling_closed_forms():
    rng = np.random.default_rng(3)
    t = rng.integers(0, 256, (3, 5, 7, 4), dtype=np.uint8)
    # texel centres return the texel exactly
    for (j, i) in [(0, 0), (4, 6), (2, 3)]:
        got = orc.tex_sample(t, (i + 0.5) / 7, (j + 0.5) / 5, 1)
        assert np.array_equal(got, t[1, j, i].astype(np.float32) / np.float32(255))
    # a constant texture is that constant everywhere
    c = np.full((1, 3, 3, 4), 200, np.uint8)
    for u, v in [(0.1, 0.9), (-3.3, 7.25), (0.5, 0.5)]:
        assert np.allclose(orc.tex_sample(c, u, v, 0), 200 / 255, rtol=0, atol=2e-7)
    # repeat addressing: u and u + 1 (both exact) agree, negative wraps
    assert np.array_equal(orc.tex_sample(t, 0.25, 0.375, 2), orc.tex_sample(t, 1.25, 1.375, 2))
    assert np.array_equal(orc.tex_sample(t, -0.75, 0.375, 2), orc.tex_sample(t, 0.25, 0.375, 2))
    # the layer is clamped: i32(TEX_NONE) = -1 reads layer 0, past the end the last
    assert np.array_equal(orc.tex_sample(t, 0.3, 0.6, 0xFFFFFFFF), orc.tex_sample(t, 0.3, 0.6, 0))
    assert np.array_equal(orc.tex_sample(t, 0.3, 0.6, 9), orc.tex_sample(t, 0.3, 0.6, 2))
    # generic points against a float64 brute-force bilinear (structure: indices, weights, wrap)
    for _ in range(200):
        u, v = rng.uniform(-2, 3, 2)
        layer = int(rng.integers(0, 3))
        got = orc.tex_sample(t, np.float32(u), np.float32(v), layer)
        want = _bilinear64(t, float(np.float32(u)), float(np.float32(v)), layer)
        assert np.allclose(got, want, rtol=0, atol=3e-6), (u, v, got, want)


def test_textured_scene_has_every_role():
    s = rt.Scene.builtin("textured")
    shapes = [s.texture_array(k).shape for k in range(4)]
    assert shapes == [(2, 64, 64, 4), (1, 32, 32, 4), (1, 16, 16, 4), (1, 32, 32, 4)]
    mats = s.materials()
    assert (mats["tex_albedo"] == 0).any() and (mats["tex_albedo"] == 1).any()
    assert ((mats["flags"] & 1) == 1).sum() == 1
    assert (mats["tex_mr"] == 0).any() and (mats["tex_emissive"] == 0).any() and (mats["tex_normal"] == 0).any()
    # the cutout layer keeps the 16x16 disc in its corner; the rest is transparent
    a = s.texture_array(0)[1]
    assert a[:16, :16, 3].max() == 255 and not a[16:, :, 3].any() and not a[:, 16:, 3].any()
