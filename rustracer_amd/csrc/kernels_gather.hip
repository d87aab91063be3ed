// kernels_gather.hip — the camera-path kernel (rustracer-renderer/shaders/gather.wgsl)
// and the fixed-ray-set closest-hit kernel, for gfx950.
//
// One lane per pixel; a 64-lane workgroup covers an 8x8 tile (one wavefront),
// keeps each lane's traversal stack in LDS, and tiles are remapped so every
// XCD works on one contiguous band of the image (better L2 reuse of the BVH
// levels a band's rays share). Per-pixel arithmetic follows gather.wgsl line
// by line (cited below), so each pixel's RNG stream and result match a CPU
// evaluation of the shader exactly.
#include "device_common.h"
#include "kernels.h"

namespace rtd {

struct GatherArgs {
    TraceScene ts;
    const rt_triangle *tris;
    const rt_material *mats;
    const rt_light *lights;
    uint32_t n_lights;
    uint32_t n_tris;
    rt_camera_uniform cam;
    rt_image_params ip;
    rt_grid_params gp;
    const uint2 *grid_meta;
    const rt_photon *sorted;
    const float4 *sorted_pos;
    float4 *frame;
    float4 *moments;
    float4 *gbuffer;
    uint32_t max_depth;
    uint32_t tiles_x;
    uint32_t n_tiles;
    unsigned long long *counters;
};

struct LightSample {
    F3 e, wi;
};

// The tracing context of one lane: closest-hit query and shadow query.
template <int MODE, bool EARLY, bool COUNT>
struct Tracer {
    const TraceScene &ts;
    LaneStack &stk;
    TraceCount cnt;
    uint32_t queries;
    __device__ Hit closest(F3 ro, F3 rd) {
        if (COUNT) queries++;
        return trace_mode<MODE, false, COUNT>(ts, ro, rd, kInf, stk, cnt);
    }
    // gather.wgsl:91-103 without alpha-cutout materials (no textures here):
    // occluded iff the closest hit is nearer than max_dist - 0.001. EARLY
    // stops the same traversal as soon as a hit that near is found.
    __device__ bool shadowed(F3 origin, F3 dir, float max_dist) {
        if (COUNT) queries++;
        const float lim = max_dist - 0.001f;
        const Hit h = trace_mode<MODE, EARLY, COUNT>(ts, origin, dir, lim, stk, cnt);
        return !(h.t >= lim);
    }
};

// gather.wgsl:107-111
__device__ __forceinline__ F3 sky_color(F3 dir, F3 zenith, F3 horizon, float intensity) {
    const float t = wclamp(dir.y * 0.5f + 0.5f, 0.0f, 1.0f);
    return mix(horizon, zenith, wgsl_pow(t, 0.5f)) * intensity;
}

// gather.wgsl:152-206 (out-of-range light index clamped to n-1)
template <class T>
__device__ LightSample direct_light(const GatherArgs &A, T &tr, F3 hp, F3 sn, Rng &rng) {
    const uint32_t n = A.n_lights;
    uint32_t li = wgsl_u32(rng_f32(rng) * (float)n);
    li = li < n ? li : n - 1u;
    const rt_light &l = A.lights[li];
    const F3 ldata = ld3(l.data), lcolor = ld3(l.color);
    if (l.kind == 0u) {
        const F3 to = ldata - hp;
        const float r2 = dot(to, to);
        const float r = wgsl_sqrt(r2);
        const F3 ld = to / r;
        if (dot(sn, ld) <= 0.0f) return LightSample{fs(0.0f), ld};
        if (tr.shadowed(hp + sn * 0.01f, ld, r - 0.001f)) return LightSample{fs(0.0f), ld};
        const F3 ple = ld3(A.ip.point_light_emission);
        return LightSample{lcolor * l.intensity * ple * wmax(dot(sn, ld), 0.0f) / (r2 + 0.01f) * (float)n, ld};
    } else if (l.kind == 1u) {
        const F3 ld = normalize(ldata);
        const float ndl = dot(sn, ld);
        if (ndl <= 0.0f) return LightSample{fs(0.0f), ld};
        if (tr.shadowed(hp + sn * 0.01f, ld, 1e30f)) return LightSample{fs(0.0f), ld};
        return LightSample{lcolor * l.intensity * ndl * (float)n, ld};
    } else if (l.kind == 2u) {
        uint32_t ti = wgsl_u32(l.data[0]);
        ti = ti < A.n_tris ? ti : A.n_tris - 1u;
        const rt_triangle &tri = A.tris[ti];
        const float u1 = rng_f32(rng);
        const float u2 = rng_f32(rng);
        const float su = wgsl_sqrt(u1);
        const float u = 1.0f - su;
        const float v = u2 * su;
        const float w = 1.0f - u - v;
        const F3 v0 = ld3(tri.v0), v1 = ld3(tri.v1), v2 = ld3(tri.v2);
        const F3 lp = v0 * w + v1 * u + v2 * v;
        const F3 le = lcolor * ld3(A.ip.light_emission);
        const F3 to = lp - hp;
        const float r2 = dot(to, to);
        const float r = wgsl_sqrt(r2);
        const F3 ld = to / r;
        const float ndl = dot(sn, ld);
        if (ndl <= 0.0f) return LightSample{fs(0.0f), ld};
        const F3 ln = normalize(cross(v1 - v0, v2 - v0));
        const float ndl_l = dot(ln, -ld);
        if (ndl_l <= 0.0f) return LightSample{fs(0.0f), ld};
        if (tr.shadowed(hp + sn * 0.01f, ld, r - 0.001f)) return LightSample{fs(0.0f), ld};
        const float area = 0.5f * length(cross(v1 - v0, v2 - v0));
        return LightSample{le * l.intensity * ndl * ndl_l * area * (float)n / (r2 + 0.01f), ld};
    }
    return LightSample{fs(0.0f), sn};
}

// gather.wgsl:210-227
__device__ __forceinline__ float smith_g1(float nd, float a) {
    const float k = a * 0.5f;
    return nd / (nd * (1.0f - k) + k);
}
__device__ __forceinline__ F3 spec_ggx(F3 wo, F3 n, F3 wi, float a, F3 f0) {
    const float ndl = dot(n, wi);
    const float ndv = dot(n, wo);
    if (ndl <= 0.0f || ndv <= 0.0f) return fs(0.0f);
    const F3 h = normalize(wo + wi);
    const float ndh = wmax(dot(n, h), 0.0f);
    const float a2 = a * a;
    const float dd = ndh * ndh * (a2 - 1.0f) + 1.0f;
    const float D = a2 / (kPi * dd * dd);
    const F3 F = fresnel_schlick(wmax(dot(wo, h), 0.0f), f0);
    return F * (D * smith_g1(ndl, a) * smith_g1(ndv, a) / (4.0f * ndl * ndv));
}

// gather.wgsl:230-256, including its final select: the mirror direction is
// returned when the sampled direction lies above the surface.
__device__ F3 sample_spec_dir(F3 rd, F3 n, float a, Rng &rng) {
    const F3 up = (__builtin_fabsf(n.y) < 0.999f) ? f3(0.0f, 1.0f, 0.0f) : f3(1.0f, 0.0f, 0.0f);
    const F3 bt = normalize(cross(up, n));
    const F3 bb = cross(n, bt);
    const F3 mrd = -rd;
    const F3 v = f3(dot(bt, mrd), dot(bb, mrd), dot(n, mrd));
    const F3 vh = normalize(f3(a * v.x, a * v.y, wmax(v.z, 1e-4f)));
    const float l2 = vh.x * vh.x + vh.y * vh.y;
    const F3 t1 = (l2 > 0.0f) ? f3(-vh.y, vh.x, 0.0f) * wgsl_inverse_sqrt(l2) : f3(1.0f, 0.0f, 0.0f);
    const F3 t2 = cross(vh, t1);
    const float r = wgsl_sqrt(rng_f32(rng));
    const float phi = (2.0f * kPi) * rng_f32(rng);
    const float x = r * wgsl_cos(phi);
    const float y = r * wgsl_sin(phi);
    const float s = 0.5f * (1.0f + vh.z);
    const float yy = (1.0f - s) * wgsl_sqrt(wmax(0.0f, 1.0f - x * x)) + s * y;
    const F3 hh = x * t1 + yy * t2 + wgsl_sqrt(wmax(0.0f, 1.0f - x * x - yy * yy)) * vh;
    const F3 hn = normalize(hh);
    const F3 m = bt * hn.x + bb * hn.y + n * hn.z;
    const F3 out = reflect(rd, m);
    return (dot(out, n) > 0.0f) ? reflect(rd, n) : out;
}

__device__ __forceinline__ float luma_of(F3 c) { return dot(c, f3(0.2126f, 0.7152f, 0.0722f)); }

// Per-pixel accumulators of gather.wgsl:36-42. The per-sample luma sums
// (sample_lum[s] += l) build up in `cur` while sample s runs (the same additions
// in the same order from 0) and are shifted into sl[] when the sample ends, so
// after n samples they sit in order at sl[8-n .. 7]. Only static indices are
// used: the array stays in registers.
struct Acc {
    F3 col;
    float lum_sum, lum2_sum;
    float cur;
    float sl[8];
    __device__ __forceinline__ void add(F3 c) {
        col = col + c;
        const float l = luma_of(c);
        lum_sum += l;
        lum2_sum += l * l;
        cur += l;
    }
    __device__ __forceinline__ void end_sample() {
#pragma unroll
        for (int k = 0; k < 7; k++) sl[k] = sl[k + 1];
        sl[7] = cur;
        cur = 0.0f;
    }
};

__device__ __forceinline__ int32_t iclamp(int32_t v, int32_t lo, int32_t hi) { return min(max(v, lo), hi); }

// The per-view part of ImageParams/CameraUniform: one camera at one resolution.
struct ViewParams {
    rt_camera_uniform cam;
    uint32_t width, height, spp, frame;
};

// gather.wgsl:258-594 for the pixel (px, py) of view `V`; the result goes to
// element `out_idx` of the frame / moments / G-buffer arrays.
template <int MODE, bool EARLY, bool COUNT>
__device__ void gather_pixel(const GatherArgs &A, const ViewParams &V, Tracer<MODE, EARLY, COUNT> &tr, uint32_t px,
                             uint32_t py, size_t out_idx, uint32_t &photon_reads) {
    const rt_image_params &ip = A.ip;
    const rt_camera_uniform &cam = V.cam;
    const uint32_t base_seed = py * V.width + px + V.frame * 2654435761u;
    Rng rng = rng_init(base_seed * 1664525u + 1013904223u);
    Acc acc;
    acc.col = fs(0.0f);
    acc.lum_sum = 0.0f;
    acc.lum2_sum = 0.0f;
    acc.cur = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; i++) acc.sl[i] = 0.0f;
    F3 first_normal = f3(0.0f, 1.0f, 0.0f);
    float first_depth = 1e30f;
    const uint32_t spp = max(1u, V.spp);
    const uint32_t debug_view = ip._pad[0];

    float env_intensity = 0.0f;
    F3 env_color = f3(0.55f, 0.6f, 0.75f);
    F3 env_horizon = f3(0.95f, 0.93f, 0.9f);
    for (uint32_t i = 0; i < A.n_lights; i++) {
        if (A.lights[i].kind == 3u) {
            env_intensity = A.lights[i].intensity;
            env_color = ld3(A.lights[i].color);
            env_horizon = ld3(A.lights[i].data);
            break;
        }
    }
    const F3 cam_pos = ld3(cam.position), cam_fwd = ld3(cam.forward);
    const F3 cam_right = ld3(cam.right), cam_up = ld3(cam.up);

    for (uint32_t s = 0u; s < spp; s++) {
        const float jx = rng_f32(rng) - 0.5f;
        const float jy = rng_f32(rng) - 0.5f;
        const float ux = ((float)px + jx) * cam.inv_half_res[0] - 1.0f;
        const float uy = 1.0f - ((float)py + jy) * cam.inv_half_res[1];
        F3 ro = cam_pos;
        F3 rd = normalize(cam_fwd + cam_right * ux + cam_up * uy);
        F3 thr = fs(1.0f);

        for (uint32_t b = 0u; b < A.max_depth; b++) {
            const Hit hit = tr.closest(ro, rd);
            if (hit.t >= kInf) {
                if (b > 0u) acc.add(thr * sky_color(rd, env_color, env_horizon, env_intensity));
                break;
            }
            const rt_triangle &trg = A.tris[hit.prim];
            const rt_material &mat = A.mats[trg.material_id];
            const F3 hp = ro + rd * hit.t;
            const Bary bb = barycentric(trg, ro, rd, hit.t);
            const F3 gn = get_triangle_normal(trg);
            const bool ff = dot(rd, gn) < 0.0f;
            const F3 sn_geom = smooth_normal(trg, bb, gn);
            const F3 sn = ff ? sn_geom : -sn_geom;
            const F3 shn = sn;

            if (b == 0u && hit.t < first_depth) {
                first_normal = sn;
                first_depth = hit.t;
            }
            const F3 albedo = f3(mat.albedo[0], mat.albedo[1], mat.albedo[2]);
            const float rough = mat.roughness;
            const float metal = mat.metallic;
            const F3 emissive = ld3(mat.emissive);
            acc.add(thr * emissive);

            if (debug_view != 0u) {
                bool stop = true;
                switch (debug_view) {
                    case 1u: acc.col = f3(bb.uv.x, bb.uv.y, 0.0f); break;
                    case 2u: case 3u: acc.col = sn * 0.5f + fs(0.5f); break;
                    case 4u: acc.col = gn * 0.5f + fs(0.5f); break;
                    case 5u: case 12u: acc.col = emissive; break;
                    case 8u: case 9u: case 11u: case 18u: acc.col = fs(0.0f); break;
                    case 13u: acc.col = f3((float)mat.tex_emissive / 65535.0f, 0.0f, 0.0f); break;
                    case 17u: acc.col = f3(bb.uv.x - floorf(bb.uv.x), bb.uv.y - floorf(bb.uv.y), 0.0f); break;
                    case 19u: acc.col = tangent_frame(trg, sn).t * 0.5f + fs(0.5f); break;
                    case 20u: acc.col = tangent_frame(trg, sn).b * 0.5f + fs(0.5f); break;
                    case 21u: acc.col = hp; break;
                    default: stop = false; break;
                }
                if (stop) break;
            }

            if (mat.kind == 2u) {
                const F3 f0 = fs(wgsl_pow((mat.ior - 1.0f) / (mat.ior + 1.0f), 2.0f));
                const float ct = __builtin_fabsf(dot(rd, shn));
                const float fr = fresnel_schlick(ct, f0).x;
                if (rng_f32(rng) < fr) {
                    rd = reflect(rd, shn);
                    ro = hp + shn * 0.001f;
                } else {
                    const float eta = ff ? 1.0f / mat.ior : mat.ior;
                    rd = refract(rd, shn, eta);
                    ro = hp - shn * 0.001f;
                    thr = thr * albedo;
                }
                continue;
            }

            const F3 f0 = mix(fs(0.04f), albedo, fs(metal));
            const F3 wo = -rd;
            const float a = wmax(rough * rough, 1e-3f);
            const F3 fr = fresnel_schlick(__builtin_fabsf(dot(rd, shn)), f0);
            const F3 diffuse_weight = (1.0f - metal) * albedo;
            const float sp = vmax3(fr);
            const float dp = vmax3(diffuse_weight);
            const float sel = (sp + dp <= 0.0f) ? 0.5f : sp / (sp + dp);
            const float spec_prob = wclamp(sel, 0.001f, 0.999f);
            const bool do_spec = rng_f32(rng) < spec_prob;
            const F3 dscale = diffuse_weight / (1.0f - spec_prob);

            const LightSample ls = direct_light(A, tr, hp, shn, rng);
            if (!do_spec) {
                if (env_intensity > 0.0f)
                    acc.add(thr * dscale * sky_color(shn, env_color, env_horizon, env_intensity));
                acc.add(thr * dscale * ls.e / kPi);
                acc.add(thr * albedo * ip.ambient);
            } else {
                acc.add(thr * spec_ggx(wo, shn, ls.wi, a, f0) * ls.e / spec_prob);
            }

            if (ip.use_photons == 1u && !do_spec) {
                // gather.wgsl:473-514
                const rt_grid_params &gp = A.gp;
                const float r2 = ip.photon_radius * ip.photon_radius;
                F3 ph_sum = fs(0.0f);
                uint32_t ph_hits = 0u, budget = 384u;
                bool done = false;
                const int32_t bx = wgsl_i32((hp.x - gp.room_min_x) / gp.cell_size);
                const int32_t by = wgsl_i32((hp.y - gp.room_min_y) / gp.cell_size);
                const int32_t bz = wgsl_i32((hp.z - gp.room_min_z) / gp.cell_size);
                const int32_t gxmax = (int32_t)gp.grid_x - 1, gymax = (int32_t)gp.grid_y - 1,
                              gzmax = (int32_t)gp.grid_z - 1;
                for (int32_t dz = -1; dz <= 1 && !done; dz++) {
                    for (int32_t dy = -1; dy <= 1 && !done; dy++) {
                        for (int32_t dx = -1; dx <= 1 && !done; dx++) {
                            const int32_t cx = iclamp((int32_t)((uint32_t)bx + (uint32_t)dx), 0, gxmax);
                            const int32_t cy = iclamp((int32_t)((uint32_t)by + (uint32_t)dy), 0, gymax);
                            const int32_t cz = iclamp((int32_t)((uint32_t)bz + (uint32_t)dz), 0, gzmax);
                            const uint32_t cell = (uint32_t)cx + (uint32_t)cy * gp.grid_x + (uint32_t)cz * gp.grid_x * gp.grid_y;
                            const uint2 meta = A.grid_meta[cell];
                            for (uint32_t i = 0u; i < meta.y && !done; i++) {
                                if (budget == 0u) {
                                    done = true;
                                    break;
                                }
                                budget -= 1u;
                                const float4 pp4 = A.sorted_pos[meta.x + i];
                                if (COUNT) photon_reads++;
                                const F3 d = xyz(pp4) - hp;
                                const float d2 = dot(d, d);
                                if (d2 < r2) {
                                    const rt_photon &ph = A.sorted[meta.x + i];
                                    const float4 pw = *reinterpret_cast<const float4 *>(&ph.power[0]);
                                    const float4 inc = *reinterpret_cast<const float4 *>(&ph.incident[0]);
                                    ph_sum = ph_sum + xyz(pw) * wmax(dot(shn, -xyz(inc)), 0.0f);
                                    ph_hits += 1u;
                                    if (ph_hits >= 64u) {
                                        done = true;
                                        break;
                                    }
                                }
                            }
                        }
                    }
                }
                if (ph_hits > 0u)
                    acc.add(thr * dscale * ph_sum * ip.photon_scale /
                                   (kPi * (float)ph_hits * ip.photon_radius * ip.photon_radius));
                if (ip.photon_debug == 1u) {
                    const float density = wmin((float)ph_hits / 64.0f, 1.0f);
                    acc.col = f3(1.0f - density, density, 0.0f) * 2.0f;
                }
            }

            if (do_spec) {
                rd = sample_spec_dir(rd, shn, a, rng);
                ro = hp + shn * 0.001f;
                const F3 h = normalize(wo + rd);
                const float ndl = wmax(dot(shn, rd), 0.0f);
                thr = thr * fresnel_schlick(wmax(dot(wo, h), 0.0f), f0) * smith_g1(ndl, a) / spec_prob;
            } else {
                const F3 local = sample_cosine_hemisphere(rng);
                rd = to_world(local, shn);
                ro = hp + shn * 0.001f;
                thr = thr * dscale;
                if (ip.use_photons == 1u) break;
            }
            const float p = vmax3(thr);
            if (p <= 0.0f || rng_f32(rng) > p) break;
            thr = thr / p;
        }
        acc.end_sample();
    }

    // gather.wgsl:548-579 firefly sigma clamp (spp <= 8). The samples' lumas are
    // at sl[w0 .. 7] (w0 = 8 - spp) in sample order; the insertion sort, median,
    // spread and clamped sums run over that window with static indices.
    if (spp <= 8u) {
        float *sl = acc.sl;
        const uint32_t w0 = 8u - spp;
#pragma unroll
        for (uint32_t i = 1u; i < 8u; i++) {
            if (i > w0) {
                const float key = sl[i];
                bool moving = true;
#pragma unroll
                for (uint32_t j = i; j > 0u; j--) {
                    if (moving) {
                        if (j - 1u >= w0 && sl[j - 1u] > key) {
                            sl[j] = sl[j - 1u];
                            if (j - 1u == w0) {
                                sl[j - 1u] = key;
                                moving = false;
                            }
                        } else {
                            sl[j] = key;
                            moving = false;
                        }
                    }
                }
            }
        }
        const uint32_t lo = w0 + (spp - 1u) / 2u, hi = w0 + spp / 2u;
        float slo = 0.0f, shi = 0.0f;
#pragma unroll
        for (uint32_t k = 0; k < 8u; k++) {
            if (k == lo) slo = sl[k];
            if (k == hi) shi = sl[k];
        }
        const float median = (slo + shi) * 0.5f;
        const uint32_t half = max(spp / 2u, 1u);
        float spread = 0.0f;
#pragma unroll
        for (uint32_t k = 0; k < 8u; k++) {
            if (k >= w0 && k < w0 + half) {
                const float d = sl[k] - median;
                spread += d * d;
            }
        }
        const float sigma = wmax(wgsl_sqrt(spread / (float)half), 0.25f * median + 1e-5f);
        const float thresh = median + 3.0f * sigma;
        float cl = 0.0f, cl2 = 0.0f;
#pragma unroll
        for (uint32_t k = 0; k < 8u; k++) {
            if (k >= w0) {
                const float lc = wmin(sl[k], thresh);
                cl += lc;
                cl2 += lc * lc;
            }
        }
        acc.col = acc.col * (cl / wmax(acc.lum_sum, 1e-10f));
        acc.lum_sum = cl;
        acc.lum2_sum = cl2;
    }
    const F3 col = vmin(acc.col / (float)spp, fs(10.0f));
    A.frame[out_idx] = make_float4(col.x, col.y, col.z, 1.0f);
    A.moments[out_idx] =
        make_float4(wmin(acc.lum_sum / (float)spp, 10.0f), wmin(acc.lum2_sum / (float)spp, 100.0f), 0.0f, 0.0f);
    A.gbuffer[out_idx] = make_float4(first_normal.x, first_normal.y, first_normal.z, first_depth);
}

template <int MODE, bool EARLY, bool COUNT>
__global__ __launch_bounds__(64) void gather_kernel(GatherArgs A) {
    __shared__ uint32_t stack_rows[kLdsStack * 64];
    // XCD-aware tile order: workgroup b runs on XCD b % 8; give each XCD a
    // contiguous band of tiles.
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t q = nb / 8u;
    const uint32_t tile = (b < 8u * q) ? (b % 8u) * q + b / 8u : b;
    if (tile >= A.n_tiles) return;
    const uint32_t px = (tile % A.tiles_x) * 8u + (threadIdx.x & 7u);
    const uint32_t py = (tile / A.tiles_x) * 8u + (threadIdx.x >> 3);
    uint32_t spill_words[kSpillAlloc];
    LaneStack stk;
    stk.spill = (priv_u32 *)spill_words;
    stk.lds = (lds_u32 *)&stack_rows[threadIdx.x];
    stk.stride = 64u;
    Tracer<MODE, EARLY, COUNT> tr{A.ts, stk, TraceCount{0u, 0u}, 0u};
    uint32_t photon_reads = 0u;
    const ViewParams V{A.cam, A.ip.width, A.ip.height, A.ip.spp, A.ip.frame};
    if (px < A.ip.width && py < A.ip.height)
        gather_pixel<MODE, EARLY, COUNT>(A, V, tr, px, py, (size_t)py * A.ip.width + px, photon_reads);
    if (COUNT) {
        atomicAdd(&A.counters[0], (unsigned long long)tr.queries);
        atomicAdd(&A.counters[1], (unsigned long long)tr.cnt.nodes);
        atomicAdd(&A.counters[2], (unsigned long long)tr.cnt.tris);
        atomicAdd(&A.counters[3], (unsigned long long)photon_reads);
    }
}

// ─────────────────────────── many views, one persistent launch ───────────────────────────

struct BatchView {
    rt_camera_uniform cam;
    uint32_t width, height, spp, frame;
    uint64_t pixel_offset;  // first element of this view in the batch outputs
    uint32_t tiles_x;
    uint32_t pad;
};

struct BatchArgs {
    GatherArgs g;             // scene, lights, global image parameters, batch output arrays
    const BatchView *views;
    const uint2 *tiles;       // (view, tile index inside the view), most expensive first
    uint32_t n_tiles;
    uint32_t tile_w, tile_h;  // tile shape; tile_w * tile_h is a multiple of 64
    uint32_t tiles_per_fetch; // queue granularity
    uint32_t *counter;        // next unclaimed tile (zeroed before the launch)
};

// Persistent workgroups (one wavefront each) pull tiles from a global queue
// until it is empty; every pixel runs the same per-pixel code as gather_kernel,
// with its own view's camera, resolution, spp and seed frame.
template <int MODE, bool EARLY>
__global__ __launch_bounds__(64) void gather_batch_kernel(BatchArgs B) {
    __shared__ uint32_t stack_rows[kLdsStack * 64];
    uint32_t spill_words[kSpillAlloc];
    LaneStack stk;
    stk.spill = (priv_u32 *)spill_words;
    stk.lds = (lds_u32 *)&stack_rows[threadIdx.x];
    stk.stride = 64u;
    Tracer<MODE, EARLY, false> tr{B.g.ts, stk, TraceCount{0u, 0u}, 0u};
    uint32_t photon_reads = 0u;
    const uint32_t rounds = (B.tile_w * B.tile_h) / 64u;
    for (;;) {
        uint32_t first = 0u;
        if (threadIdx.x == 0) first = atomicAdd(B.counter, B.tiles_per_fetch);
        first = __builtin_amdgcn_readfirstlane(__shfl(first, 0, 64));
        if (first >= B.n_tiles) break;
        const uint32_t last = min(first + B.tiles_per_fetch, B.n_tiles);
        for (uint32_t t = first; t < last; t++) {
            const uint2 te = B.tiles[t];
            const uint32_t vi = __builtin_amdgcn_readfirstlane(te.x);
            const uint32_t ti = __builtin_amdgcn_readfirstlane(te.y);
            const BatchView &bv = B.views[vi];
            const ViewParams V{bv.cam, bv.width, bv.height, bv.spp, bv.frame};
            const uint32_t tx = ti % bv.tiles_x, ty = ti / bv.tiles_x;
            for (uint32_t r = 0; r < rounds; r++) {
                const uint32_t k = r * 64u + threadIdx.x;
                const uint32_t px = tx * B.tile_w + k % B.tile_w, py = ty * B.tile_h + k / B.tile_w;
                if (px < V.width && py < V.height)
                    gather_pixel<MODE, EARLY, false>(B.g, V, tr, px, py, bv.pixel_offset + (size_t)py * V.width + px,
                                                     photon_reads);
            }
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(64) void trace_rays_kernel(TraceScene ts, const rt_ray *rays, uint32_t n, rt_hit *out) {
    __shared__ uint32_t stack_rows[kLdsStack * 64];
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= n) return;
    uint32_t spill_words[kSpillAlloc];
    LaneStack stk;
    stk.spill = (priv_u32 *)spill_words;
    stk.lds = (lds_u32 *)&stack_rows[threadIdx.x];
    stk.stride = 64u;
    TraceCount cnt{0u, 0u};
    const rt_ray r = rays[i];
    const F3 ro = ld3(r.origin), rd = ld3(r.dir);
    const Hit h = trace_mode<MODE, false, false>(ts, ro, rd, kInf, stk, cnt);
    out[i].t = h.t;
    out[i].prim_id = h.prim;
}

TraceScene make_trace_scene(const rtk::TraceInputs &t) {
    TraceScene s;
    s.nodes = t.nodes;
    s.prims = t.prims;
    s.slot_tris = reinterpret_cast<const float4 *>(t.slot_tris);
    s.wnodes = reinterpret_cast<const float4 *>(t.wnodes);
    s.w4nodes = reinterpret_cast<const float4 *>(t.w4nodes);
    s.n_nodes = t.n_nodes;
    s.pad = 0;
    s.root_min = make_float4(t.root_min[0], t.root_min[1], t.root_min[2], t.root_min[3]);
    s.root_max = make_float4(t.root_max[0], t.root_max[1], t.root_max[2], t.root_max[3]);
    return s;
}

}  // namespace rtd

namespace rtk {

static rtd::GatherArgs gather_args(const GatherLaunch &g) {
    rtd::GatherArgs A;
    A.ts = rtd::make_trace_scene(g.trace);
    A.tris = g.tris;
    A.mats = g.mats;
    A.lights = g.lights;
    A.n_lights = g.n_lights;
    A.n_tris = g.n_tris;
    A.cam = g.cam;
    A.ip = g.ip;
    A.gp = g.gp;
    A.grid_meta = reinterpret_cast<const uint2 *>(g.grid_meta);
    A.sorted = g.sorted_photons;
    A.sorted_pos = reinterpret_cast<const float4 *>(g.sorted_pos);
    A.frame = reinterpret_cast<float4 *>(g.frame);
    A.moments = reinterpret_cast<float4 *>(g.moments);
    A.gbuffer = reinterpret_cast<float4 *>(g.gbuffer);
    A.max_depth = g.max_depth;
    A.tiles_x = (g.ip.width + 7u) / 8u;
    const uint32_t tiles_y = (g.ip.height + 7u) / 8u;
    A.n_tiles = A.tiles_x * tiles_y;
    A.counters = g.counters;
    return A;
}

hipError_t launch_gather(const GatherLaunch &g, hipStream_t s) {
    const rtd::GatherArgs A = gather_args(g);
    const dim3 grid(A.n_tiles), block(64);
#define RT_GATHER(M, E, C) hipLaunchKernelGGL((rtd::gather_kernel<M, E, C>), grid, block, 0, s, A)
#define RT_GATHER_M(M)                        \
    if (g.early_shadow) {                     \
        if (g.counters) RT_GATHER(M, true, true); \
        else RT_GATHER(M, true, false);       \
    } else {                                  \
        if (g.counters) RT_GATHER(M, false, true); \
        else RT_GATHER(M, false, false);      \
    }
    if (g.mode == 0) { RT_GATHER_M(0) }
    else if (g.mode == 1) { RT_GATHER_M(1) }
    else if (g.mode == 2) { RT_GATHER_M(2) }
    else { RT_GATHER_M(3) }
#undef RT_GATHER_M
#undef RT_GATHER
    return hipGetLastError();
}

hipError_t launch_gather_batch(const GatherLaunch &g, const BatchLaunch &b, hipStream_t s) {
    static_assert(sizeof(rtd::BatchView) == sizeof(BatchViewRecord), "batch view record layout");
    if (b.n_tiles == 0) return hipSuccess;
    rtd::BatchArgs B;
    B.g = gather_args(g);
    B.views = reinterpret_cast<const rtd::BatchView *>(b.views);
    B.tiles = reinterpret_cast<const uint2 *>(b.tiles);
    B.n_tiles = b.n_tiles;
    B.tile_w = b.tile_w;
    B.tile_h = b.tile_h;
    B.tiles_per_fetch = b.tiles_per_fetch;
    B.counter = b.counter;
    hipError_t e = launch_fill_u32(b.counter, 0u, 1, s);
    if (e != hipSuccess) return e;
    const dim3 grid(b.n_workgroups), block(64);
#define RT_BATCH(M, E) hipLaunchKernelGGL((rtd::gather_batch_kernel<M, E>), grid, block, 0, s, B)
    if (g.mode == 0) { if (g.early_shadow) RT_BATCH(0, true); else RT_BATCH(0, false); }
    else if (g.mode == 1) { if (g.early_shadow) RT_BATCH(1, true); else RT_BATCH(1, false); }
    else if (g.mode == 2) { if (g.early_shadow) RT_BATCH(2, true); else RT_BATCH(2, false); }
    else { if (g.early_shadow) RT_BATCH(3, true); else RT_BATCH(3, false); }
#undef RT_BATCH
    return hipGetLastError();
}

hipError_t launch_trace_rays(const TraceInputs &t, uint32_t mode, const rt_ray *rays, uint32_t n, rt_hit *out,
                             hipStream_t s) {
    if (n == 0) return hipSuccess;
    const rtd::TraceScene ts = rtd::make_trace_scene(t);
    const dim3 grid((n + 63u) / 64u), block(64);
    if (mode == 0) hipLaunchKernelGGL((rtd::trace_rays_kernel<0>), grid, block, 0, s, ts, rays, n, out);
    else if (mode == 1) hipLaunchKernelGGL((rtd::trace_rays_kernel<1>), grid, block, 0, s, ts, rays, n, out);
    else if (mode == 2) hipLaunchKernelGGL((rtd::trace_rays_kernel<2>), grid, block, 0, s, ts, rays, n, out);
    else hipLaunchKernelGGL((rtd::trace_rays_kernel<3>), grid, block, 0, s, ts, rays, n, out);
    return hipGetLastError();
}

}  // namespace rtk
