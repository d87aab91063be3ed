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

#endif /* RT_TEXTURE_H */
