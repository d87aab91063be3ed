/*
 * rt_texture.h — material texture arrays and their sampling, shared by the
 * gfx950 kernels and the CPU oracle so both take the same value.
 *
 * A role (albedo, metallic-roughness, emissive, normal) is one 2D array of
 * rgba8 texels (lib.rs:1286-1358, create_role_array): every layer has the size
 * of the role's first texture, a smaller texture fills the top-left corner and
 * the rest stays zero, and an empty role is one 1x1 layer of (0,0,0,0). Texels
 * are the linear floats quantised as (u8)(clamp(c, 0, 1) * 255) (truncation).
 *
 * textureSampleLevel(tex, sampler, uv, i32(layer), 0.0) with the renderer's
 * sampler (repeat addressing, linear min/mag filter, level 0; lib.rs:508-517):
 * WGSL clamps the layer to [0, layers - 1] (i32(TEX_NONE) = -1 reads layer 0);
 * the filter is the Vulkan bilinear rule x = u*w - 0.5, i0 = floor(x), a = x -
 * i0, texel indices wrapped into [0, w). The fractional weights and the
 * combination are implementation-defined on the GPU; this backend fixes them as
 * float32 mix(mix(t00, t10, a), mix(t01, t11, a), b) with mix(p, q, t) =
 * p*(1-t) + q*t and unorm8 -> float as byte / 255.
 */
#ifndef RT_TEXTURE_H
#define RT_TEXTURE_H

#include <stdint.h>

#include "rt_types.h"
#include "rt_wgsl_math.h"

RT_WM_FN int64_t rt_tex_wrap(int64_t i, int64_t n) {
    const int64_t m = i % n;
    return m < 0 ? m + n : m;
}

RT_WM_FN float rt_tex_mix(float p, float q, float t) { return p * (1.0f - t) + q * t; }

/* The filtered rgba at uv of layer i32(layer_bits). */
RT_WM_FN void rt_tex_sample(const rt_tex_array *t, float u, float v, uint32_t layer_bits, float out[4]) {
    int32_t layer = (int32_t)layer_bits;
    const int32_t last = (int32_t)t->layers - 1;
    layer = layer < 0 ? 0 : (layer > last ? last : layer);
    const float x = u * (float)t->width - 0.5f;
    const float y = v * (float)t->height - 0.5f;
    const float fx = __builtin_floorf(x), fy = __builtin_floorf(y);
    const float ax = x - fx, ay = y - fy;
    const int64_t w = (int64_t)t->width, h = (int64_t)t->height;
    const int64_t x0 = rt_tex_wrap((int64_t)wgsl_i32(fx), w), x1 = rt_tex_wrap((int64_t)wgsl_i32(fx) + 1, w);
    const int64_t y0 = rt_tex_wrap((int64_t)wgsl_i32(fy), h), y1 = rt_tex_wrap((int64_t)wgsl_i32(fy) + 1, h);
    const uint8_t *base = t->texels + (uint64_t)layer * (uint64_t)(w * h) * 4u;
    const uint8_t *p00 = base + (y0 * w + x0) * 4, *p10 = base + (y0 * w + x1) * 4;
    const uint8_t *p01 = base + (y1 * w + x0) * 4, *p11 = base + (y1 * w + x1) * 4;
    for (int c = 0; c < 4; c++) {
        const float t00 = (float)p00[c] / 255.0f, t10 = (float)p10[c] / 255.0f;
        const float t01 = (float)p01[c] / 255.0f, t11 = (float)p11[c] / 255.0f;
        out[c] = rt_tex_mix(rt_tex_mix(t00, t10, ax), rt_tex_mix(t01, t11, ax), ay);
    }
}

/*
 * The present pass (fullscreen.wgsl, drawn by Renderer::record_frame,
 * lib.rs:834-850): the tonemapped rgba8 output texture sampled through the
 * quad sampler (ClampToEdge, linear min/mag filter, no mip levels;
 * lib.rs:555-562) at the fullscreen triangle's interpolated uv, written to a
 * Bgra8Unorm surface (rustracer-app/src/main.rs:320-324). The surface keeps
 * its own size when the window is resized (main.rs:156-162), so the pass
 * resamples the render-size output to the surface size.
 *
 * Surface pixel p of n_s along an axis has its fragment centre at p + 0.5 and
 * interpolated coordinate u = (p + 0.5) / n_s (fullscreen.wgsl:16-19: uv =
 * ((x + 1) / 2, (1 - y) / 2) over clip space, rows top to bottom); the Vulkan
 * bilinear rule samples texel position x = u*n_t - 0.5 between texels floor(x)
 * and floor(x) + 1, both clamped into [0, n_t). The interpolation precision,
 * the sub-texel precision of the weights and the float -> unorm8 rounding are
 * implementation-defined on the GPU; this backend fixes them as:
 *   x = (p + 0.5) * s - 0.5 in float32, with s = n_t / n_s rounded once;
 *   the fraction x - floor(x) rounded to 8 bits (round to nearest even):
 *   a = rint((x - floor(x)) * 256) in 0..256, weights 256 - a and a (the
 *   sub-texel precision Vulkan drivers report as subTexelPrecisionBits = 8);
 *   per channel the exact integer blend S = wy0*(wx0*t00 + wx1*t10) +
 *   wy1*(wx0*t01 + wx1*t11) of the bytes, and the unorm8 result
 *   (S + 32768) >> 16, the bilinear value S / 65536 rounded half up.
 */
RT_WM_FN float rt_present_scale(uint32_t n_t, uint32_t n_s) { return (float)n_t / (float)n_s; }

RT_WM_FN void rt_present_axis(uint32_t p, float scale, uint32_t n_t, uint32_t *i0, uint32_t *i1, uint32_t *a) {
    const float x = ((float)p + 0.5f) * scale - 0.5f;
    const float f = __builtin_floorf(x);
    /* 0..256; the mask changes no value and lets the GPU compiler use 24-bit multiplies */
    *a = (uint32_t)__builtin_rintf((x - f) * 256.0f) & 0x1FFu;
    const int32_t last = (int32_t)n_t - 1;
    const int32_t k = (int32_t)f; /* x >= -0.5, so k >= -1 */
    *i0 = (uint32_t)(k < 0 ? 0 : (k > last ? last : k));
    *i1 = (uint32_t)(k + 1 > last ? last : (k + 1 < 0 ? 0 : k + 1));
}

/* Four rgba8 texels (byte 0 = r) filtered with 8-bit weights (ax, ay), as one
 * bgra8 surface pixel (byte 0 = b, byte 3 = a). */
RT_WM_FN uint32_t rt_present_blend(uint32_t t00, uint32_t t10, uint32_t t01, uint32_t t11, uint32_t ax,
                                   uint32_t ay) {
    ax &= 0x1FFu;
    ay &= 0x1FFu;
    const uint32_t bx = (256u - ax) & 0x1FFu, by = (256u - ay) & 0x1FFu;
    uint32_t out = 0;
    for (int c = 0; c < 4; c++) {
        const int sh = 8 * c;
        const uint32_t r0 = bx * ((t00 >> sh) & 0xFFu) + ax * ((t10 >> sh) & 0xFFu);
        const uint32_t r1 = bx * ((t01 >> sh) & 0xFFu) + ax * ((t11 >> sh) & 0xFFu);
        const uint32_t q = (by * r0 + ay * r1 + 32768u) >> 16;
        const int dst = c == 3 ? 24 : 8 * (2 - c); /* r -> byte 2, g -> 1, b -> 0, a -> 3 */
        out |= q << dst;
    }
    return out;
}

#endif /* RT_TEXTURE_H */
