// kernels_post.hip — per-frame image passes for gfx950: the a-trous filter
// (rustracer-renderer/shaders/denoise.wgsl), composite (composite.wgsl) fused
// with tonemap (tonemap.wgsl), and the FP64 running sum used to judge convergence.
//
// All three are HBM-bound streams over float4 pixel buffers. The a-trous pass
// stages its workgroup's tile plus the 2*step halo of colour and G-buffer in
// LDS once, so the 25 taps per pixel read LDS instead of re-fetching 32 bytes
// from L2 per tap; the variance of the 3x3 moment neighbourhood is read the same way.
#include "device_common.h"
#include "kernels.h"

namespace rtd {

constexpr int kDnTile = 16;  // 16x16 output pixels per workgroup (256 lanes)

__device__ __forceinline__ float luma_of(F3 c) { return dot(c, f3(0.2126f, 0.7152f, 0.0722f)); }

// denoise.wgsl:51-124. LDS holds, for the (16 + 4*step)^2 halo tile: colour rgb
// and luminance, G-buffer normal and depth. step <= 4 fits (32^2 * 32 B = 32 KB).
template <int STEP>
__global__ __launch_bounds__(256) void denoise_kernel(const float4 *__restrict__ in, float4 *__restrict__ out,
                                                      const float4 *__restrict__ gbuf, const float4 *__restrict__ mom,
                                                      rt_denoise_params p, const rtk::FrameState *fstate) {
    if (fstate) p.alpha = fstate->alpha;
    constexpr int H = 2 * STEP;             // halo radius of the 5x5 taps
    constexpr int T = kDnTile + 2 * H;      // staged tile side
    __shared__ float4 s_col[T * T];         // rgb + luminance
    __shared__ float4 s_gb[T * T];          // normal + depth
    __shared__ float s_var[(kDnTile + 2) * (kDnTile + 2)];  // raw per-pixel variance, 1-pixel halo
    const int W = (int)p.width, Hh = (int)p.height;
    const int x0 = (int)blockIdx.x * kDnTile, y0 = (int)blockIdx.y * kDnTile;
    for (int i = (int)threadIdx.x; i < T * T; i += 256) {
        const int gx = x0 - H + i % T, gy = y0 - H + i / T;
        float4 c = make_float4(0.f, 0.f, 0.f, 0.f), g = make_float4(0.f, 0.f, 0.f, 0.f);
        if (gx >= 0 && gy >= 0 && gx < W && gy < Hh) {
            const size_t idx = (size_t)gy * p.width + (uint32_t)gx;
            c = in[idx];
            g = gbuf[idx];
            c.w = luma_of(xyz(c));
        }
        s_col[i] = c;
        s_gb[i] = g;
    }
    constexpr int V = kDnTile + 2;
    for (int i = (int)threadIdx.x; i < V * V; i += 256) {
        const int gx = x0 - 1 + i % V, gy = y0 - 1 + i / V;
        float v = 0.0f;
        if (gx >= 0 && gy >= 0 && gx < W && gy < Hh) {
            const float4 m = mom[(size_t)gy * p.width + (uint32_t)gx];
            v = wmax(m.y - m.x * m.x, 0.0f);
        }
        s_var[i] = v;
    }
    __syncthreads();
    const int lx = (int)threadIdx.x % kDnTile, ly = (int)threadIdx.x / kDnTile;
    const int px = x0 + lx, py = y0 + ly;
    if (px >= W || py >= Hh) return;
    const float4 cc = s_col[(ly + H) * T + (lx + H)];
    const float4 gc = s_gb[(ly + H) * T + (lx + H)];
    const F3 center = xyz(cc);
    const F3 n_center = xyz(gc);
    const float d_center = gc.w;
    const float lum_center = cc.w;
    // denoise.wgsl:70-85, accumulated in the reference's (qy, qx) order
    float var_acc = 0.0f, var_cnt = 0.0f;
    for (int qy = -1; qy <= 1; qy++) {
        for (int qx = -1; qx <= 1; qx++) {
            const int nx = px + qx, ny = py + qy;
            if (nx < 0 || ny < 0 || nx >= W || ny >= Hh) continue;
            var_acc += s_var[(ly + 1 + qy) * V + (lx + 1 + qx)];
            var_cnt += 1.0f;
        }
    }
    const float var_frame = var_acc / wmax(var_cnt, 1.0f);
    const float ema_scale = p.alpha / wmax(2.0f - p.alpha, 1e-4f);
    const float var_c = var_frame * ema_scale;
    const float kern[3] = {0.375f, 0.25f, 0.0625f};
    F3 sum = fs(0.0f);
    float total_weight = 0.0f;
#pragma unroll
    for (int dy = -2; dy <= 2; dy++) {
#pragma unroll
        for (int dx = -2; dx <= 2; dx++) {
            const int nx = px + dx * STEP, ny = py + dy * STEP;
            if (nx < 0 || ny < 0 || nx >= W || ny >= Hh) continue;
            const int li = (ly + H + dy * STEP) * T + (lx + H + dx * STEP);
            const float4 nc = s_col[li];
            const float4 ng = s_gb[li];
            const float sw = kern[dx < 0 ? -dx : dx] * kern[dy < 0 ? -dy : dy];
            const float nw = wgsl_pow(wmax(dot(n_center, xyz(ng)), 0.0f), p.phi_normal);
            const float dw = wgsl_exp(-__builtin_fabsf(d_center - ng.w) / (p.phi_depth * wmax(d_center, 1e-4f)));
            const float dl = nc.w - lum_center;
            const float cw = wgsl_exp(-(dl * dl) / (var_c * p.lambda + 1e-6f));
            const float w = sw * nw * dw * cw;
            sum = sum + xyz(nc) * w;
            total_weight += w;
        }
    }
    F3 o = center;
    if (total_weight > 1e-6f) o = sum / total_weight;
    out[(size_t)py * p.width + px] = make_float4(o.x, o.y, o.z, 1.0f);
}

// Generic-step fallback (steps > 4): the same arithmetic reading global memory.
__global__ __launch_bounds__(256) void denoise_global_kernel(const float4 *__restrict__ in, float4 *__restrict__ out,
                                                             const float4 *__restrict__ gbuf,
                                                             const float4 *__restrict__ mom, rt_denoise_params p,
                                                             const rtk::FrameState *fstate) {
    if (fstate) p.alpha = fstate->alpha;
    const int W = (int)p.width, Hh = (int)p.height;
    const int px = (int)(blockIdx.x * kDnTile + threadIdx.x % kDnTile);
    const int py = (int)(blockIdx.y * kDnTile + threadIdx.x / kDnTile);
    if (px >= W || py >= Hh) return;
    const int step = (int)p.step_size;
    const size_t idx = (size_t)py * p.width + px;
    const F3 center = xyz(in[idx]);
    const float4 gc = gbuf[idx];
    const F3 n_center = xyz(gc);
    const float d_center = gc.w;
    const float lum_center = luma_of(center);
    float var_acc = 0.0f, var_cnt = 0.0f;
    for (int qy = -1; qy <= 1; qy++) {
        for (int qx = -1; qx <= 1; qx++) {
            const int nx = px + qx, ny = py + qy;
            if (nx < 0 || ny < 0 || nx >= W || ny >= Hh) continue;
            const float4 m = mom[(size_t)ny * p.width + nx];
            var_acc += wmax(m.y - m.x * m.x, 0.0f);
            var_cnt += 1.0f;
        }
    }
    const float var_c = (var_acc / wmax(var_cnt, 1.0f)) * (p.alpha / wmax(2.0f - p.alpha, 1e-4f));
    const float kern[3] = {0.375f, 0.25f, 0.0625f};
    F3 sum = fs(0.0f);
    float total_weight = 0.0f;
    for (int dy = -2; dy <= 2; dy++) {
        for (int dx = -2; dx <= 2; dx++) {
            const int nx = px + dx * step, ny = py + dy * step;
            if (nx < 0 || ny < 0 || nx >= W || ny >= Hh) continue;
            const size_t n = (size_t)ny * p.width + nx;
            const F3 nb = xyz(in[n]);
            const float4 ng = gbuf[n];
            const float sw = kern[dx < 0 ? -dx : dx] * kern[dy < 0 ? -dy : dy];
            const float nw = wgsl_pow(wmax(dot(n_center, xyz(ng)), 0.0f), p.phi_normal);
            const float dw = wgsl_exp(-__builtin_fabsf(d_center - ng.w) / (p.phi_depth * wmax(d_center, 1e-4f)));
            const float dl = luma_of(nb) - lum_center;
            const float cw = wgsl_exp(-(dl * dl) / (var_c * p.lambda + 1e-6f));
            const float w = sw * nw * dw * cw;
            sum = sum + nb * w;
            total_weight += w;
        }
    }
    F3 o = center;
    if (total_weight > 1e-6f) o = sum / total_weight;
    out[idx] = make_float4(o.x, o.y, o.z, 1.0f);
}

// tonemap.wgsl:19-50 for one pixel and the rgba8unorm store (clamp to [0,1],
// scale by 255, round half up).
__device__ __forceinline__ uint32_t tonemap_pixel(float4 a, uint32_t i, uint32_t width, const rt_tonemap_uniform &tm,
                                                  const float4 *__restrict__ gbuf) {
    const F3 exposed = xyz(a) * tm.exposure;
    const F3 mapped = exposed / (exposed + fs(1.0f));
    const float inv_gamma = 1.0f / tm.gamma;
    F3 color = f3(wgsl_pow(mapped.x, inv_gamma), wgsl_pow(mapped.y, inv_gamma), wgsl_pow(mapped.z, inv_gamma));
    if (tm.marker[3] > 0.0f) {
        const uint32_t x = i % width, y = i / width;
        const float dxp = ((float)x + 0.5f) - tm.marker[0];
        const float dyp = ((float)y + 0.5f) - tm.marker[1];
        const float d = wgsl_sqrt(dxp * dxp + dyp * dyp);
        const float r = wmax(tm.marker[2], 1.5f);
        if (gbuf[i].w > tm.marker[3] - 0.2f) {
            const float ball = 1.0f - wgsl_smoothstep(r * 0.75f, r, d);
            const float halo = 1.0f - wgsl_smoothstep(r, r * 3.5f, d);
            const F3 core = f3(1.0f, 0.94f, 0.80f) * wgsl_pow(ball, 1.4f) * 0.85f;
            const F3 ring = f3(1.0f, 0.46f, 0.14f) * halo * halo * 0.8f;
            color = color + (core + ring) * tm.glow * (fs(1.0f) - 0.55f * color);
        }
    }
    const float c4[3] = {color.x, color.y, color.z};
    uint32_t packed = 0xFF000000u;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float v = wclamp(c4[k], 0.0f, 1.0f);
        packed |= ((uint32_t)floorf(v * 255.0f + 0.5f) & 0xFFu) << (8 * k);
    }
    return packed;
}

// composite.wgsl:16-26 then the tonemap of the new accumulation. One pass over the pixels.
__global__ __launch_bounds__(256) void composite_tonemap_kernel(const float4 *__restrict__ frame,
                                                                float4 *__restrict__ accum, uint32_t *__restrict__ rgba8,
                                                                uint32_t n_px, uint32_t width, float alpha,
                                                                rt_tonemap_uniform tm, const float4 *__restrict__ gbuf,
                                                                const rtk::FrameState *fstate) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n_px) return;
    if (fstate) alpha = fstate->alpha;
    const float4 fresh = frame[i];
    const float4 prev = accum[i];
    const float om = 1.0f - alpha;
    const float4 a = make_float4(prev.x * om + fresh.x * alpha, prev.y * om + fresh.y * alpha,
                                 prev.z * om + fresh.z * alpha, prev.w * om + fresh.w * alpha);
    accum[i] = a;
    rgba8[i] = tonemap_pixel(a, i, width, tm, gbuf);
}

// Running luminance moments beside the accumulation, weighted as it is.
__global__ __launch_bounds__(256) void moment_history_kernel(const float4 *__restrict__ moments,
                                                             float2 *__restrict__ hist, uint32_t n_px, float alpha,
                                                             const rtk::FrameState *fstate) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n_px) return;
    if (fstate) alpha = fstate->alpha;
    const float4 m = moments[i];
    const float2 h = hist[i];
    const float om = 1.0f - alpha;
    hist[i] = make_float2(h.x * om + m.x * alpha, h.y * om + m.y * alpha);
}

// One lane stores a frame graph's per-frame values (vector store).
__global__ void set_frame_state_kernel(rtk::FrameState *dst, rtk::FrameState v) {
    if (threadIdx.x == 0) *dst = v;
}

// The tonemap alone, of another displayed image (the learned denoiser's output).
__global__ __launch_bounds__(256) void tonemap_kernel(const float4 *__restrict__ img, uint32_t *__restrict__ rgba8,
                                                      uint32_t n_px, uint32_t width, rt_tonemap_uniform tm,
                                                      const float4 *__restrict__ gbuf) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n_px) return;
    rgba8[i] = tonemap_pixel(img[i], i, width, tm, gbuf);
}

__global__ __launch_bounds__(256) void accum64_add_kernel(const float4 *__restrict__ acc, double *__restrict__ sum,
                                                          size_t n_px) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= n_px) return;
    const float4 a = acc[i];
    sum[4 * i + 0] += (double)a.x;
    sum[4 * i + 1] += (double)a.y;
    sum[4 * i + 2] += (double)a.z;
    sum[4 * i + 3] += (double)a.w;
}

// relMSE of an image against the FP64 mean sum/n: sum over pixels and rgb of
// (x - y)^2 / (y^2 + eps), in double. One partial per block (fixed grid-stride
// order and a fixed tree inside the block), then one block adds the partials
// in index order: the result does not depend on scheduling.
constexpr int kRelBlocks = 1024;
__global__ __launch_bounds__(256) void relmse_partial_kernel(const float4 *__restrict__ img,
                                                             const double *__restrict__ sum, double inv_n, double eps,
                                                             size_t n_px, double *__restrict__ partial) {
    __shared__ double red[256];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < n_px; i += (size_t)gridDim.x * 256u) {
        const float4 x = img[i];
        const double xs[3] = {(double)x.x, (double)x.y, (double)x.z};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double y = sum[4 * i + c] * inv_n;
            const double d = xs[c] - y;
            acc += d * d / (y * y + eps);
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(256) void relmse_final_kernel(const double *__restrict__ partial, int n, double scale,
                                                           double *__restrict__ out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * scale;
}

// Plain streaming copy, one 16-byte load and store per lane and no loop: the
// measured-HBM denominator of the rooflines and the calibration kernel of the
// traffic counters. Measured on a 1 GiB copy: 6.2 TB/s (read + write) for this
// form; grid-stride loops over 2k-32k workgroups with one or four loads in
// flight per lane reached 4.2-5.1 TB/s, hipMemcpy device-to-device 5.0 TB/s.
__global__ __launch_bounds__(256) void hbm_copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                       size_t n) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// The present pass (fullscreen.wgsl): a wavefront per 256 pixels of a surface
// row, four per lane at a stride of 64, so every store instruction writes 256
// contiguous bytes and the row's vertical taps and weights are worked out once
// per lane. The four texels of each pixel come from two output rows that
// neighbouring lanes share (L2 after the first touch). The blend is integer
// (24-bit multiplies on bytes picked out by SDWA operands); with the float
// sample positions and 64-bit addresses it is about 77 VALU instructions per
// pixel, which bounds the pass before HBM does (1080p -> 4K: 0.022 ms).
constexpr uint32_t kPresentPx = 4;  // pixels per lane

__global__ __launch_bounds__(256) void present_kernel(const uint32_t *__restrict__ rgba8, uint32_t tw, uint32_t th,
                                                      uint32_t *__restrict__ surface, uint32_t sw, uint32_t sh,
                                                      uint32_t pitch_words, float scale_x, float scale_y) {
    const uint32_t y = blockIdx.y * 4u + (threadIdx.x >> 6);
    if (y >= sh) return;
    uint32_t y0, y1, ay;
    rt_present_axis(y, scale_y, th, &y0, &y1, &ay);
    const uint32_t *r0 = rgba8 + (size_t)y0 * tw, *r1 = rgba8 + (size_t)y1 * tw;
    uint32_t *out = surface + (size_t)y * pitch_words;
    const uint32_t xb = blockIdx.x * (64u * kPresentPx) + (threadIdx.x & 63u);
#pragma unroll
    for (uint32_t k = 0; k < kPresentPx; k++) {
        const uint32_t x = xb + 64u * k;
        if (x < sw) {
            uint32_t x0, x1, ax;
            rt_present_axis(x, scale_x, tw, &x0, &x1, &ax);
            out[x] = rt_present_blend(r0[x0], r0[x1], r1[x0], r1[x1], ax, ay);
        }
    }
}

}  // namespace rtd

namespace rtk {

hipError_t launch_tonemap(const void *img, uint32_t *rgba8, uint32_t w, uint32_t h, const rt_tonemap_uniform &tm,
                          const void *gbuffer, hipStream_t s) {
    const uint32_t n = w * h;
    hipLaunchKernelGGL(rtd::tonemap_kernel, dim3((n + 255u) / 256u), dim3(256), 0, s,
                       reinterpret_cast<const float4 *>(img), rgba8, n, w, tm,
                       reinterpret_cast<const float4 *>(gbuffer));
    return hipGetLastError();
}

size_t relmse_scratch_doubles() { return rtd::kRelBlocks + 1; }

hipError_t launch_relmse(const void *img, const double *sum, uint64_t n_frames, float eps, size_t n_px,
                         double *scratch, hipStream_t s) {
    const double inv_n = n_frames ? 1.0 / (double)n_frames : 0.0;
    hipLaunchKernelGGL(rtd::relmse_partial_kernel, dim3(rtd::kRelBlocks), dim3(256), 0, s,
                       reinterpret_cast<const float4 *>(img), sum, inv_n, (double)eps, n_px, scratch);
    hipLaunchKernelGGL(rtd::relmse_final_kernel, dim3(1), dim3(256), 0, s, scratch, rtd::kRelBlocks,
                       1.0 / (3.0 * (double)n_px), scratch + rtd::kRelBlocks);
    return hipGetLastError();
}

hipError_t launch_hbm_copy(const void *src, void *dst, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 16;
    if (n == 0) return hipSuccess;
    const size_t blocks = (n + 255) / 256;
    if (blocks > 0x7FFFFFFFu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(rtd::hbm_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<const float4 *>(src), reinterpret_cast<float4 *>(dst), n);
    return hipGetLastError();
}

hipError_t launch_denoise(const DenoiseLaunch &d, hipStream_t s) {
    const dim3 grid((d.p.width + rtd::kDnTile - 1) / rtd::kDnTile, (d.p.height + rtd::kDnTile - 1) / rtd::kDnTile);
    const float4 *in = reinterpret_cast<const float4 *>(d.in);
    float4 *out = reinterpret_cast<float4 *>(d.out);
    const float4 *gb = reinterpret_cast<const float4 *>(d.gbuffer);
    const float4 *mo = reinterpret_cast<const float4 *>(d.moments);
    switch (d.p.step_size) {
        case 1: hipLaunchKernelGGL((rtd::denoise_kernel<1>), grid, dim3(256), 0, s, in, out, gb, mo, d.p, d.fstate); break;
        case 2: hipLaunchKernelGGL((rtd::denoise_kernel<2>), grid, dim3(256), 0, s, in, out, gb, mo, d.p, d.fstate); break;
        case 4: hipLaunchKernelGGL((rtd::denoise_kernel<4>), grid, dim3(256), 0, s, in, out, gb, mo, d.p, d.fstate); break;
        default: hipLaunchKernelGGL(rtd::denoise_global_kernel, grid, dim3(256), 0, s, in, out, gb, mo, d.p, d.fstate); break;
    }
    return hipGetLastError();
}

hipError_t launch_composite_tonemap(const void *frame, void *accum, uint32_t *rgba8, uint32_t w, uint32_t h,
                                    float alpha, const rt_tonemap_uniform &tm, const void *gbuffer,
                                    const FrameState *fstate, hipStream_t s) {
    const uint32_t n = w * h;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(rtd::composite_tonemap_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                       reinterpret_cast<const float4 *>(frame), reinterpret_cast<float4 *>(accum), rgba8, n, w, alpha,
                       tm, reinterpret_cast<const float4 *>(gbuffer), fstate);
    return hipGetLastError();
}

hipError_t launch_moment_history(const void *moments, void *hist, uint32_t n_px, float alpha,
                                 const FrameState *fstate, hipStream_t s) {
    if (n_px == 0) return hipSuccess;
    hipLaunchKernelGGL(rtd::moment_history_kernel, dim3((n_px + 255) / 256), dim3(256), 0, s,
                       reinterpret_cast<const float4 *>(moments), reinterpret_cast<float2 *>(hist), n_px, alpha,
                       fstate);
    return hipGetLastError();
}

hipError_t launch_set_frame_state(FrameState *dst, const FrameState &v, hipStream_t s) {
    hipLaunchKernelGGL(rtd::set_frame_state_kernel, dim3(1), dim3(64), 0, s, dst, v);
    return hipGetLastError();
}

hipError_t launch_present(const uint32_t *rgba8, uint32_t tw, uint32_t th, uint32_t *surface, uint32_t sw,
                          uint32_t sh, uint32_t pitch_words, hipStream_t s) {
    if (sw == 0 || sh == 0) return hipSuccess;
    const uint32_t per_block = 64u * rtd::kPresentPx;
    hipLaunchKernelGGL(rtd::present_kernel, dim3((sw + per_block - 1u) / per_block, (sh + 3u) / 4u), dim3(256), 0, s,
                       rgba8, tw, th, surface, sw, sh, pitch_words, rt_present_scale(tw, sw), rt_present_scale(th, sh));
    return hipGetLastError();
}

hipError_t launch_accum64_add(const void *accum_f4, double *sum, size_t n_px, hipStream_t s) {
    if (n_px == 0) return hipSuccess;
    hipLaunchKernelGGL(rtd::accum64_add_kernel, dim3((unsigned)((n_px + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const float4 *>(accum_f4), sum, n_px);
    return hipGetLastError();
}

}  // namespace rtk
