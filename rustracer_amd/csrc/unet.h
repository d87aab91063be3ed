// unet.h — the learned denoiser's host-side entry used by the renderer.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "../../include/rustracer_hip.h"

namespace rtu {

// The renderer's device buffers a denoising pass reads.
struct FrameViews {
    const void *accum;    // displayed linear image before denoising (float4 per pixel)
    const void *gbuf;     // normal.xyz, depth (float4)
    const void *albedo;   // first-hit albedo of the pixel-centre ray (float4)
    const void *moments;  // accumulated luminance moments E[l], E[l^2] (float2)
    uint32_t width, height;
    uint32_t frames;      // frames accumulated so far
    float ema;            // alpha / (2 - alpha) of the accumulation's current weight
};

// The renderer's current buffers, with the albedo pass brought up to date and
// the renderer's work finished (implemented by the renderer).
FrameViews frame_views(rt_renderer *r);

// Runs the network on `v` and writes the denoised float4 image to `out`
// (width * height float4), on `stream`.
void denoise(rt_unet *u, const FrameViews &v, void *out, hipStream_t stream);

}  // namespace rtu
