// kernels_gather.hip — the camera-path kernel (rustracer-renderer/shaders/gather.wgsl)
// and the fixed-ray-set closest-hit kernel, for gfx950.
//
// One lane per pixel; a 64-lane workgroup covers an 8x8 tile (one wavefront),
// keeps each lane's traversal stack in LDS, and tiles are remapped so every
// XCD works on one contiguous band of the image (better L2 reuse of the BVH
// levels a band's rays share). Per-pixel arithmetic follows gather.wgsl line
// by line (cited below), so each pixel's RNG stream and result match a CPU
// evaluation of the shader exactly.
#include <algorithm>

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
    const float4 *gather_rec;  // 3 per sorted photon: (pos.xyz, power.x), (power.yz, incident.xy), (incident.z)
    float4 *frame;
    float4 *moments;
    float4 *gbuffer;
    uint32_t max_depth;
    uint32_t tiles_x;
    uint32_t n_tiles;
    unsigned long long *counters;
    // path-queue (wavefront) state; unused by the one-pass kernels
    float4 *px_state;       // kPixelPlanes planes of n_px float4
    float4 *queue[2];       // kQueuePlanes planes of n_px float4 each: queue[b & 1] holds the paths entering depth b
    uint32_t *path_count;   // per depth b >= 1 at 3 (b - 1): work taken, diffuse paths queued, smooth paths queued
    uint32_t n_px;
    uint32_t sample;        // sample index of this pass
    rt_tex_array tex[RT_TEX_ROLES];  // albedo, metallic-roughness, emissive, normal
    uint32_t *regen;        // persistent kernel: kBands pixel counters, zero at launch
    uint32_t refill_min;    // persistent kernel: lanes without a pixel before a wavefront takes new ones
    const rtk::FrameState *fstate;  // non-null: camera and seed frame of a replayed frame graph
    uint32_t n_frames;      // persistent kernel: consecutive frames of this launch (>= 1)
    uint32_t frame_stride;  // float4 elements between the outputs of consecutive frames
};

// The direct light's sample at a shading point: the unoccluded contribution,
// the direction, and whether the shader casts a shadow ray (and its reach).
struct LightSample {
    F3 e, wi;
    float reach;
    bool cast;
};

// The tracing context of one lane: closest-hit query and shadow query.
template <int MODE, bool EARLY, bool COUNT>
struct Tracer {
    const TraceScene &ts;
    LaneStack &stk;
    TraceCount cnt;
    __device__ Hit closest(F3 ro, F3 rd) {
        const uint32_t t0 = COUNT ? phase_clock() : 0u;
        if (COUNT) {
            cnt.add(kCntQueries, 1);
            cnt.trav_phase = kPhaseTraversal;
        }
        const Hit h = trace_mode<MODE, false, COUNT>(ts, ro, rd, kInf, stk, cnt);
        if (COUNT) count_time(cnt, kPhaseTraversal, t0);
        return h;
    }
    // gather.wgsl:92-93: is the closest hit nearer than max_dist - 0.001? EARLY
    // stops the same traversal as soon as a hit that near is found (then `h`
    // is that hit, not necessarily the closest).
    __device__ bool shadowed(F3 origin, F3 dir, float max_dist, Hit &h) {
        if (COUNT) {
            cnt.add(kCntQueries, 1);
            cnt.trav_phase = kPhaseShadow;
        }
        const uint32_t t0 = COUNT ? phase_clock() : 0u;
        const float lim = max_dist - 0.001f;
        h = trace_mode<MODE, EARLY, COUNT>(ts, origin, dir, lim, stk, cnt);
        if (COUNT) count_time(cnt, kPhaseShadow, t0);
        return !(h.t >= lim);
    }
};

// gather.wgsl:91-103 with alpha-cutout: the closest hit decides, and a
// transparent texel of a cutout material lets the shadow ray through. Scenes
// with cutout materials run with EARLY = false (the closest hit is needed).
// `h` is the answer of the query (the closest hit, or with EARLY the first hit
// found nearer than lim = max_dist - 0.001).
__device__ bool shadow_blocked(const GatherArgs &A, const Hit &h, F3 origin, F3 dir, float lim) {
    if (h.t >= lim) return false;
    const rt_triangle &bt = A.tris[h.prim];
    const rt_material &bm = A.mats[bt.material_id];
    if ((bm.flags & RT_MAT_FLAG_ALPHA_CUTOUT) != 0u && bm.tex_albedo != RT_TEX_NONE) {
        const F2 uv = barycentric(bt, origin, dir, h.t).uv;
        if (tex_sample(A.tex[RT_TEX_ROLE_ALBEDO], uv, bm.tex_albedo).w < 0.5f) return false;
    }
    return true;
}
template <class T>
__device__ bool occluded(const GatherArgs &A, T &tr, F3 origin, F3 dir, float max_dist) {
    Hit h;
    tr.shadowed(origin, dir, max_dist, h);
    return shadow_blocked(A, h, origin, dir, max_dist - 0.001f);
}

// gather.wgsl:119-141
struct MatParams {
    F3 albedo;
    float rough, metal;
};
__device__ __forceinline__ MatParams eval_material(const GatherArgs &A, const rt_material &mat, F2 uv) {
    MatParams p{f3(mat.albedo[0], mat.albedo[1], mat.albedo[2]), mat.roughness, mat.metallic};
    if (mat.tex_albedo != RT_TEX_NONE) p.albedo = p.albedo * rgb(tex_sample(A.tex[RT_TEX_ROLE_ALBEDO], uv, mat.tex_albedo));
    if (mat.tex_mr != RT_TEX_NONE) {
        const F4s mr = tex_sample(A.tex[RT_TEX_ROLE_MR], uv, mat.tex_mr);
        p.rough = wclamp(mat.roughness * mr.y, 0.0f, 1.0f);
        p.metal = wclamp(mat.metallic * mr.z, 0.0f, 1.0f);
    }
    return p;
}
__device__ __forceinline__ F3 emissive_at(const GatherArgs &A, const rt_material &mat, F2 uv) {
    if (mat.tex_emissive != RT_TEX_NONE)
        return ld3(mat.emissive) * rgb(tex_sample(A.tex[RT_TEX_ROLE_EMISSIVE], uv, mat.tex_emissive));
    return ld3(mat.emissive);
}

// gather.wgsl:107-111
__device__ __forceinline__ F3 sky_color(F3 dir, F3 zenith, F3 horizon, float intensity) {
    const float t = wclamp(dir.y * 0.5f + 0.5f, 0.0f, 1.0f);
    return mix(horizon, zenith, wgsl_pow(t, 0.5f)) * intensity;
}

// gather.wgsl:152-206 up to the shadow query (out-of-range light index
// clamped to n-1). Each light kind sets the direction, the shadow ray's reach
// and the unoccluded contribution; one shadow query after the branches serves
// every kind, so the lanes of a wavefront that picked different kinds
// traverse together. The arithmetic of each kind is the shader's, in its order.
__device__ LightSample sample_light(const GatherArgs &A, F3 hp, F3 sn, Rng &rng) {
    const uint32_t n = A.n_lights;
    uint32_t li = wgsl_u32(rng_f32(rng) * (float)n);
    li = li < n ? li : n - 1u;
    const rt_light &l = A.lights[li];
    const F3 ldata = ld3(l.data), lcolor = ld3(l.color);
    F3 ld = sn, e = fs(0.0f);
    float reach = 0.0f;
    bool cast = false;
    if (l.kind == 0u) {
        const F3 to = ldata - hp;
        const float r2 = dot(to, to);
        const float r = wgsl_sqrt(r2);
        ld = to / r;
        if (!(dot(sn, ld) <= 0.0f)) {  // the shader's `<= 0 -> return`, NaN included
            cast = true;
            reach = r - 0.001f;
            const F3 ple = ld3(A.ip.point_light_emission);
            e = lcolor * l.intensity * ple * wmax(dot(sn, ld), 0.0f) / (r2 + 0.01f) * (float)n;
        }
    } else if (l.kind == 1u) {
        ld = normalize(ldata);
        const float ndl = dot(sn, ld);
        if (!(ndl <= 0.0f)) {
            cast = true;
            reach = 1e30f;
            e = lcolor * l.intensity * ndl * (float)n;
        }
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
        // gather.wgsl:180-189: the emissive texture at the sampled point tints the light
        const rt_material &lmat = A.mats[tri.material_id];
        const F2 luv = F2{(tri.uv0[0] * w + tri.uv1[0] * u) + tri.uv2[0] * v,
                          (tri.uv0[1] * w + tri.uv1[1] * u) + tri.uv2[1] * v};
        F3 le = lcolor;
        if (lmat.tex_emissive != RT_TEX_NONE) le = le * rgb(tex_sample(A.tex[RT_TEX_ROLE_EMISSIVE], luv, lmat.tex_emissive));
        le = le * ld3(A.ip.light_emission);
        const F3 to = lp - hp;
        const float r2 = dot(to, to);
        const float r = wgsl_sqrt(r2);
        ld = to / r;
        const float ndl = dot(sn, ld);
        if (!(ndl <= 0.0f)) {
            const F3 ln = normalize(cross(v1 - v0, v2 - v0));
            const float ndl_l = dot(ln, -ld);
            if (!(ndl_l <= 0.0f)) {
                cast = true;
                reach = r - 0.001f;
                const float area = 0.5f * length(cross(v1 - v0, v2 - v0));
                e = le * l.intensity * ndl * ndl_l * area * (float)n / (r2 + 0.01f);
            }
        }
    }
    return LightSample{e, ld, reach, cast};
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
// after n samples they sit in order at sl[NS-n .. NS-1]. Only static indices
// are used: the array stays in registers. NS = 8 holds every sample the clamp
// looks at (spp <= 8); NS = 1 serves 1-spp passes with one register instead
// of eight.
template <int NS>
struct AccT {
    F3 col;
    float lum_sum, lum2_sum;
    float cur;
    float sl[NS];
    __device__ __forceinline__ void add(F3 c) {
        col = col + c;
        const float l = luma_of(c);
        lum_sum += l;
        lum2_sum += l * l;
        cur += l;
    }
    __device__ __forceinline__ void end_sample() {
#pragma unroll
        for (int k = 0; k < NS - 1; k++) sl[k] = sl[k + 1];
        sl[NS - 1] = cur;
        cur = 0.0f;
    }
};
using Acc = AccT<8>;

__device__ __forceinline__ int32_t iclamp(int32_t v, int32_t lo, int32_t hi) { return min(max(v, lo), hi); }

// The per-view part of ImageParams/CameraUniform: one camera at one resolution.
struct ViewParams {
    rt_camera_uniform cam;
    uint32_t width, height, spp, frame;
};

// The frame's camera and seed: from the arguments, or from the device-resident
// frame state of a replayed frame graph.
__device__ __forceinline__ ViewParams frame_view(const GatherArgs &A) {
    if (A.fstate) return ViewParams{A.fstate->cam, A.ip.width, A.ip.height, A.ip.spp, A.fstate->frame};
    return ViewParams{A.cam, A.ip.width, A.ip.height, A.ip.spp, A.ip.frame};
}

// The environment light (gather.wgsl:274-288): the first kind-3 light, else
// the shader's default sky.
struct Env {
    float intensity;
    F3 color, horizon;
};
__device__ __forceinline__ Env find_env(const GatherArgs &A) {
    Env env{0.0f, f3(0.55f, 0.6f, 0.75f), f3(0.95f, 0.93f, 0.9f)};
    for (uint32_t i = 0; i < A.n_lights; i++) {
        if (A.lights[i].kind == 3u) {
            env.intensity = A.lights[i].intensity;
            env.color = ld3(A.lights[i].color);
            env.horizon = ld3(A.lights[i].data);
            break;
        }
    }
    return env;
}

// Normal and depth of the nearest first hit over the pixel's samples (G-buffer).
struct FirstHit {
    F3 normal;
    float depth;
};

#ifndef RT_PHOTON_BATCH
#define RT_PHOTON_BATCH 4
#endif
// photon candidates whose gather records are loaded together in the radius gather
constexpr uint32_t kPhotonBatch = RT_PHOTON_BATCH;

// One iteration of the bounce loop of gather.wgsl:298-537 at depth b, in two
// parts around the direct light's shadow query.
//
// bounce_hit: the shading of the closest hit `hit` of the ray (ro, rd) up to
// the light sample. kPathEnds: the path ends (sky, debug view); kPathGoesOn:
// ro, rd and thr hold the next ray (dielectric, or a cutout texel passed);
// kSurfaceLit: `sf` holds the surface and its light sample, and the shadow
// query of sf.ls (when sf.ls.cast) must be answered before bounce_light.
//
// bounce_light: the rest of the bounce once the shadow query is answered:
// direct light, photon radius gather, next direction and Russian roulette.
// Returns true when the path goes on (ro, rd and thr then hold the next ray).
//
// When `smooth` is given it is set to whether the next ray was scattered by a
// specular lobe or a dielectric (rather than diffusely, or passed through a
// cutout texel).
enum BounceStep : uint32_t { kPathEnds = 0u, kPathGoesOn = 1u, kSurfaceLit = 2u };
struct SurfaceState {
    F3 hp, shn;                 // hit point, shading normal
    F3 f0, dscale, albedo;
    float a, spec_prob;
    bool do_spec;
    LightSample ls;
};

template <class AccType>
__device__ uint32_t bounce_hit(const GatherArgs &A, const Env &env, const Hit &hit, uint32_t b, F3 &ro, F3 &rd,
                               F3 &thr, Rng &rng, AccType &acc, FirstHit &fh, uint32_t px, uint32_t py, bool *smooth,
                               SurfaceState &sf) {
    if (smooth) *smooth = false;
    const rt_image_params &ip = A.ip;
    const uint32_t debug_view = ip._pad[0];
    if (hit.t >= kInf) {
        if (b > 0u) acc.add(thr * sky_color(rd, env.color, env.horizon, env.intensity));
        return kPathEnds;
    }
    const rt_triangle &trg = A.tris[hit.prim];
    const rt_material &mat = A.mats[trg.material_id];
    const F3 hp = ro + rd * hit.t;
    const Bary bb = barycentric(trg, ro, rd, hit.t);
    // gather.wgsl:315-323: a transparent texel of a cutout material is skipped
    // (the path goes on from just behind it; the bounce count still advances)
    if ((mat.flags & RT_MAT_FLAG_ALPHA_CUTOUT) != 0u && mat.tex_albedo != RT_TEX_NONE) {
        if (tex_sample(A.tex[RT_TEX_ROLE_ALBEDO], bb.uv, mat.tex_albedo).w < 0.5f) {
            ro = hp + rd * 0.002f;
            return kPathGoesOn;
        }
    }
    const F3 gn = get_triangle_normal(trg);
    const bool ff = dot(rd, gn) < 0.0f;
    const F3 sn_geom = smooth_normal(trg, bb, gn);
    const F3 sn = ff ? sn_geom : -sn_geom;
    F3 shn = sn;
    if (mat.tex_normal != RT_TEX_NONE) {  // gather.wgsl:334-340
        const F3 nm = rgb(tex_sample(A.tex[RT_TEX_ROLE_NORMAL], bb.uv, mat.tex_normal)) * 2.0f - fs(1.0f);
        const TanFrame frm = tangent_frame(trg, sn);
        shn = normalize(frm.t * nm.x + frm.b * nm.y + sn * nm.z);
    }

    if (b == 0u && hit.t < fh.depth) {
        fh.normal = sn;
        fh.depth = hit.t;
    }
    const MatParams params = eval_material(A, mat, bb.uv);
    const F3 albedo = params.albedo;
    const float rough = params.rough;
    const float metal = params.metal;
    const F3 emissive = emissive_at(A, mat, bb.uv);
    acc.add(thr * emissive);

    if (debug_view != 0u) {
        bool stop = true;
        switch (debug_view) {
            case 1u: acc.col = f3(bb.uv.x, bb.uv.y, 0.0f); break;
            case 2u: acc.col = sn * 0.5f + fs(0.5f); break;
            case 3u: acc.col = shn * 0.5f + fs(0.5f); break;
            case 4u: acc.col = gn * 0.5f + fs(0.5f); break;
            case 5u: acc.col = emissive; break;
            case 12u: acc.col = ld3(mat.emissive); break;
            case 8u: case 9u: {  // screen-space samples of layer 0 (gather.wgsl:365-376)
                const F2 puv = F2{(float)px / (float)ip.width, (float)py / (float)ip.height};
                acc.col = rgb(tex_sample(A.tex[debug_view == 8u ? RT_TEX_ROLE_EMISSIVE : RT_TEX_ROLE_ALBEDO], puv, 0u));
                break;
            }
            case 11u: acc.col = rgb(tex_sample(A.tex[RT_TEX_ROLE_EMISSIVE], bb.uv, 0u)); break;
            case 18u: acc.col = rgb(tex_sample(A.tex[RT_TEX_ROLE_ALBEDO], bb.uv, mat.tex_albedo)); break;
            case 13u: acc.col = f3((float)mat.tex_emissive / 65535.0f, 0.0f, 0.0f); break;
            case 17u: acc.col = f3(bb.uv.x - floorf(bb.uv.x), bb.uv.y - floorf(bb.uv.y), 0.0f); break;
            case 19u: acc.col = tangent_frame(trg, sn).t * 0.5f + fs(0.5f); break;
            case 20u: acc.col = tangent_frame(trg, sn).b * 0.5f + fs(0.5f); break;
            case 21u: acc.col = hp; break;
            default: stop = false; break;
        }
        if (stop) return kPathEnds;
    }

    if (mat.kind == 2u) {
        if (smooth) *smooth = true;
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
        return kPathGoesOn;
    }

    const F3 f0 = mix(fs(0.04f), albedo, fs(metal));
    const float a = wmax(rough * rough, 1e-3f);
    const F3 fr = fresnel_schlick(__builtin_fabsf(dot(rd, shn)), f0);
    const F3 diffuse_weight = (1.0f - metal) * albedo;
    const float sp = vmax3(fr);
    const float dp = vmax3(diffuse_weight);
    const float sel = (sp + dp <= 0.0f) ? 0.5f : sp / (sp + dp);
    const float spec_prob = wclamp(sel, 0.001f, 0.999f);
    const bool do_spec = rng_f32(rng) < spec_prob;
    const F3 dscale = diffuse_weight / (1.0f - spec_prob);
    sf.hp = hp;
    sf.shn = shn;
    sf.f0 = f0;
    sf.dscale = dscale;
    sf.albedo = albedo;
    sf.a = a;
    sf.spec_prob = spec_prob;
    sf.do_spec = do_spec;
    sf.ls = sample_light(A, hp, shn, rng);
    return kSurfaceLit;
}

template <bool COUNT, class AccType>
__device__ bool bounce_light(const GatherArgs &A, const Env &env, const SurfaceState &sf, bool occl, F3 &ro, F3 &rd,
                             F3 &thr, Rng &rng, AccType &acc, TraceCount &cnt, bool *smooth) {
    const rt_image_params &ip = A.ip;
    const F3 hp = sf.hp, shn = sf.shn, f0 = sf.f0, dscale = sf.dscale, albedo = sf.albedo;
    const float a = sf.a, spec_prob = sf.spec_prob;
    const bool do_spec = sf.do_spec;
    const F3 wo = -rd;
    LightSample ls = sf.ls;
    if (occl) ls.e = fs(0.0f);
    if (!do_spec) {
        if (env.intensity > 0.0f)
            acc.add(thr * dscale * sky_color(shn, env.color, env.horizon, env.intensity));
        acc.add(thr * dscale * ls.e / kPi);
        acc.add(thr * albedo * ip.ambient);
    } else {
        acc.add(thr * spec_ggx(wo, shn, ls.wi, a, f0) * ls.e / spec_prob);
    }

    const uint32_t t_ph = COUNT ? phase_clock() : 0u;
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
        // Same visiting order, budget and stop rules as the shader's loops: rows
        // (dz, dy) in order, each row's three cells (one contiguous run away
        // from the x borders: grid_meta is an exclusive scan), candidates in
        // order. Each lane walks its own cursor over that sequence and takes its
        // next batch of up to kPhotonBatch candidates every step, so the
        // wavefront runs as many steps as its longest walk rather than, row by
        // row, as many as that row's longest. The next row's cell records are
        // loaded while the current row is walked; a batch's gather records are
        // loaded together. Sums are added in candidate order.
        const bool interior = bx >= 1 && bx <= gxmax - 1;
        auto load_row = [&](uint32_t r, uint2 m[3]) {
            const int32_t dz = (int32_t)(r / 3u) - 1, dy = (int32_t)(r % 3u) - 1;
            const int32_t cy = iclamp((int32_t)((uint32_t)by + (uint32_t)dy), 0, gymax);
            const int32_t cz = iclamp((int32_t)((uint32_t)bz + (uint32_t)dz), 0, gzmax);
            const uint32_t row = (uint32_t)cy * gp.grid_x + (uint32_t)cz * gp.grid_x * gp.grid_y;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int32_t cx = iclamp((int32_t)((uint32_t)bx + (uint32_t)(k - 1)), 0, gxmax);
                m[k] = A.grid_meta[(uint32_t)cx + row];
            }
        };
        uint2 nxt[3], rows[3];
        load_row(0u, nxt);
        uint32_t next_row = 0u;  // the row whose records nxt holds
        int n_ranges = 0, k_range = 0;
        uint32_t cur = 0u, left = 0u;
        while (!done) {
            // the next non-empty range of the walk, opening rows as needed
            while (left == 0u) {
                if (k_range < n_ranges) {
                    cur = rows[k_range].x;
                    left = rows[k_range].y;
                    k_range++;
                    continue;
                }
                if (next_row == 9u) break;
#pragma unroll
                for (int k = 0; k < 3; k++) rows[k] = nxt[k];
                if (interior) rows[0].y = rows[0].y + rows[1].y + rows[2].y;
                n_ranges = interior ? 1 : 3;
                k_range = 0;
                next_row++;
                if (next_row < 9u) load_row(next_row, nxt);
            }
            if (left == 0u) break;  // all 27 cells walked
            if (COUNT) count_step(cnt, kPhasePhotons);
            // one round trip per batch: position, power and incident direction
            // of each candidate come in one 48-byte record
            const uint32_t nb = min(left, kPhotonBatch);
            float4 g0[kPhotonBatch], g1[kPhotonBatch];
            float g2[kPhotonBatch];
#pragma unroll
            for (uint32_t j = 0; j < kPhotonBatch; j++) {
                if (j < nb) {
                    const float4 *gr = A.gather_rec + 3u * (cur + j);
                    g0[j] = gr[0];
                    g1[j] = gr[1];
                    g2[j] = gr[2].x;
                } else {
                    g0[j] = g1[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    g2[j] = 0.0f;
                }
            }
#pragma unroll
            for (uint32_t j = 0; j < kPhotonBatch; j++) {
                if (j < nb && !done) {
                    if (budget == 0u) {
                        done = true;
                    } else {
                        budget -= 1u;
                        if (COUNT) cnt.add(kCntPhotons, 1);
                        const F3 d = xyz(g0[j]) - hp;
                        const float d2 = dot(d, d);
                        if (d2 < r2) {
                            const F3 pw = f3(g0[j].w, g1[j].x, g1[j].y);
                            const F3 inc = f3(g1[j].z, g1[j].w, g2[j]);
                            ph_sum = ph_sum + pw * wmax(dot(shn, -inc), 0.0f);
                            ph_hits += 1u;
                            if (ph_hits >= 64u) done = true;
                        }
                    }
                }
            }
            cur += nb;
            left -= nb;
        }
        if (ph_hits > 0u)
            acc.add(thr * dscale * ph_sum * ip.photon_scale /
                           (kPi * (float)ph_hits * ip.photon_radius * ip.photon_radius));
        if (ip.photon_debug == 1u) {
            const float density = wmin((float)ph_hits / 64.0f, 1.0f);
            acc.col = f3(1.0f - density, density, 0.0f) * 2.0f;
        }
    }
    if (COUNT && ip.use_photons == 1u) count_time(cnt, kPhasePhotons, t_ph);

    if (do_spec) {
        if (smooth) *smooth = true;
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
        if (ip.use_photons == 1u) return false;
    }
    const float p = vmax3(thr);
    if (p <= 0.0f || rng_f32(rng) > p) return false;
    thr = thr / p;
    return true;
}

// One whole bounce for callers that trace in place: closest hit, shading, the
// shadow query, the rest of the shading.
template <int MODE, bool EARLY, bool COUNT, class AccType>
__device__ bool bounce(const GatherArgs &A, const Env &env, Tracer<MODE, EARLY, COUNT> &tr, uint32_t b, F3 &ro,
                       F3 &rd, F3 &thr, Rng &rng, AccType &acc, FirstHit &fh, uint32_t px,
                       uint32_t py, bool *smooth = nullptr) {
    const uint32_t t0 = COUNT ? phase_clock() : 0u;
    if (COUNT) count_step(tr.cnt, kPhaseBounce);
    const Hit hit = tr.closest(ro, rd);
    SurfaceState sf;
    const uint32_t st = bounce_hit(A, env, hit, b, ro, rd, thr, rng, acc, fh, px, py, smooth, sf);
    bool cont = st == kPathGoesOn;
    if (st == kSurfaceLit) {
        const bool occl = sf.ls.cast && occluded(A, tr, sf.hp + sf.shn * 0.01f, sf.ls.wi, sf.ls.reach);
        cont = bounce_light<COUNT>(A, env, sf, occl, ro, rd, thr, rng, acc, tr.cnt, smooth);
    }
    if (COUNT) count_time(tr.cnt, kPhaseBounce, t0);
    return cont;
}

// gather.wgsl:262-263: the pixel's seed; gather.wgsl:291-296: the jittered camera ray of one sample.
__device__ __forceinline__ Rng pixel_rng(const ViewParams &V, uint32_t px, uint32_t py, uint32_t frame) {
    const uint32_t base_seed = py * V.width + px + frame * 2654435761u;
    return rng_init(base_seed * 1664525u + 1013904223u);
}
__device__ __forceinline__ Rng pixel_rng(const ViewParams &V, uint32_t px, uint32_t py) {
    return pixel_rng(V, px, py, V.frame);
}
__device__ __forceinline__ void camera_ray(const ViewParams &V, uint32_t px, uint32_t py, Rng &rng, F3 &ro, F3 &rd) {
    const rt_camera_uniform &cam = V.cam;
    const float jx = rng_f32(rng) - 0.5f;
    const float jy = rng_f32(rng) - 0.5f;
    const float ux = ((float)px + jx) * cam.inv_half_res[0] - 1.0f;
    const float uy = 1.0f - ((float)py + jy) * cam.inv_half_res[1];
    ro = ld3(cam.position);
    rd = normalize(ld3(cam.forward) + ld3(cam.right) * ux + ld3(cam.up) * uy);
}

template <int NS>
__device__ __forceinline__ void acc_clear(AccT<NS> &acc, FirstHit &fh) {
    acc.col = fs(0.0f);
    acc.lum_sum = 0.0f;
    acc.lum2_sum = 0.0f;
    acc.cur = 0.0f;
#pragma unroll
    for (int i = 0; i < NS; i++) acc.sl[i] = 0.0f;
    fh.normal = f3(0.0f, 1.0f, 0.0f);
    fh.depth = 1e30f;
}

// gather.wgsl:540-595: firefly clamp over the samples, then the frame, moments
// and G-buffer writes.
template <int NS>
__device__ void finalize(const GatherArgs &A, AccT<NS> &acc, uint32_t spp, const FirstHit &fh, size_t out_idx) {
    // gather.wgsl:548-579 firefly sigma clamp (spp <= 8). The samples' lumas are
    // at sl[w0 .. NS-1] (w0 = NS - spp; spp <= NS whenever spp <= 8) in sample
    // order; the insertion sort, median, spread and clamped sums run over that
    // window with static indices.
    if (spp <= 8u) {
        float *sl = acc.sl;
        const uint32_t w0 = (uint32_t)NS - spp;
#pragma unroll
        for (uint32_t i = 1u; i < (uint32_t)NS; i++) {
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
        for (uint32_t k = 0; k < (uint32_t)NS; k++) {
            if (k == lo) slo = sl[k];
            if (k == hi) shi = sl[k];
        }
        const float median = (slo + shi) * 0.5f;
        const uint32_t half = max(spp / 2u, 1u);
        float spread = 0.0f;
#pragma unroll
        for (uint32_t k = 0; k < (uint32_t)NS; k++) {
            if (k >= w0 && k < w0 + half) {
                const float d = sl[k] - median;
                spread += d * d;
            }
        }
        const float sigma = wmax(wgsl_sqrt(spread / (float)half), 0.25f * median + 1e-5f);
        const float thresh = median + 3.0f * sigma;
        float cl = 0.0f, cl2 = 0.0f;
#pragma unroll
        for (uint32_t k = 0; k < (uint32_t)NS; k++) {
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
    A.gbuffer[out_idx] = make_float4(fh.normal.x, fh.normal.y, fh.normal.z, fh.depth);
}

// gather.wgsl:258-594 for the pixel (px, py) of view `V` in one pass (all
// samples and bounces); the result goes to element `out_idx` of the frame /
// moments / G-buffer arrays.
template <int MODE, bool EARLY, bool COUNT>
__device__ void gather_pixel(const GatherArgs &A, const ViewParams &V, Tracer<MODE, EARLY, COUNT> &tr, uint32_t px,
                             uint32_t py, size_t out_idx) {
    Rng rng = pixel_rng(V, px, py);
    Acc acc;
    FirstHit fh;
    acc_clear(acc, fh);
    const uint32_t spp = max(1u, V.spp);
    const Env env = find_env(A);
    for (uint32_t s = 0u; s < spp; s++) {
        F3 ro, rd;
        camera_ray(V, px, py, rng, ro, rd);
        F3 thr = fs(1.0f);
        for (uint32_t b = 0u; b < A.max_depth; b++)
            if (!bounce<MODE, EARLY, COUNT>(A, env, tr, b, ro, rd, thr, rng, acc, fh, px, py)) break;
        acc.end_sample();
    }
    finalize(A, acc, spp, fh, out_idx);
}

// At least 4 wavefronts per SIMD (a 128-VGPR budget): measured faster than the
// compiler's own choice of 130 VGPRs at 3 wavefronts.
template <int MODE, bool EARLY, bool COUNT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void gather_kernel(GatherArgs A) {
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
    stk.init(stack_rows, spill_words);
    __shared__ unsigned long long cnt_words[kCntWords];
    if (COUNT) count_begin(cnt_words);
    Tracer<MODE, EARLY, COUNT> tr{A.ts, stk, TraceCount{(lds_u64 *)cnt_words, 0u}};
    const ViewParams V = frame_view(A);
    if (px < A.ip.width && py < A.ip.height)
        gather_pixel<MODE, EARLY, COUNT>(A, V, tr, px, py, (size_t)py * A.ip.width + px);
    if (COUNT) {
        count_flush(A.counters, cnt_words);
    }
}

// ─────────────────────────── persistent lanes, pixels handed out ───────────────────────────
//
// As many one-wavefront workgroups as fit on the GPU at once. Each lane owns
// one pixel at a time and runs its samples one bounce per loop iteration; a
// lane whose pixel is done takes the next one from a shared counter, so the
// few long glass/metal paths of a frame no longer hold wavefronts of finished
// pixels resident. The pixels (in 8x8 tile order) are cut into kBands
// contiguous bands, one per XCD (workgroup b runs on XCD b % 8): a wavefront
// takes pixels from its own XCD's band first, then from the following bands,
// so neighbouring pixels share an L2. Every pixel runs exactly the steps of
// gather_pixel (same RNG stream, same accumulation order); only the lane that
// runs it and the moment it starts change.
constexpr uint32_t kBands = 8;
#ifndef RT_PERSISTENT_WAVES
#define RT_PERSISTENT_WAVES 4  // wavefronts per SIMD the register budget is cut for
#endif

// The pixel hand-out of the persistent forms: a wavefront's lanes without a
// pixel take the next slots of its current band; a slot past the image (edge
// tiles) is skipped by taking another. A band's counter runs over the band's
// slots of frame 0, then of frame 1, ... of the frames of the launch.
struct PixelFeed {
    uint32_t band, bands_left, n_slots, band_len;
    __device__ explicit PixelFeed(const GatherArgs &A)
        : band(blockIdx.x % kBands),
          bands_left(kBands),
          n_slots(A.n_tiles * 64u),
          band_len((A.n_tiles + kBands - 1u) / kBands * 64u) {}
    // Lanes with live == false take a pixel (px, py of frame fi) when there
    // is one; returns true on the lanes that took one. A wavefront takes
    // pixels once A.refill_min of its lanes have none (or none has one).
    __device__ bool take(const GatherArgs &A, bool &live, uint32_t &px, uint32_t &py, uint32_t &fi) {
        const uint32_t lane = __lane_id();
        const unsigned long long below = (1ull << lane) - 1ull;
        bool took = false;
        while (bands_left > 0u) {
            const unsigned long long need = __ballot(!live);
            if (need == 0ull) break;
            if ((uint32_t)__popcll(need) < A.refill_min && __ballot(live) != 0ull) break;
            const uint32_t n_need = (uint32_t)__popcll(need);
            const int leader = __ffsll((long long)need) - 1;
            uint32_t base = 0u;
            if (lane == (uint32_t)leader) base = atomicAdd(&A.regen[band], n_need);
            base = __builtin_amdgcn_readfirstlane(__shfl(base, leader, 64));
            const uint32_t band0 = band * band_len;
            const uint32_t len = band0 < n_slots ? min(band_len, n_slots - band0) : 0u;
            const uint32_t total = len * A.n_frames;
            if (!live) {
                const uint32_t k = base + (uint32_t)__popcll(need & below);
                if (k < total) {
                    const uint32_t f = k / len;
                    const uint32_t slot = band0 + (k - f * len), tile = slot >> 6, l = slot & 63u;
                    px = (tile % A.tiles_x) * 8u + (l & 7u);
                    py = (tile / A.tiles_x) * 8u + (l >> 3);
                    if (px < A.ip.width && py < A.ip.height) {
                        live = true;
                        took = true;
                        fi = f;
                    }
                }
            }
            if (base + n_need >= total) {
                band = (band + 1u) % kBands;
                bands_left--;
            }
        }
        return took;
    }
};

template <int MODE, bool EARLY, bool COUNT, int NS = 8>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RT_PERSISTENT_WAVES))) void gather_persistent_kernel(
    GatherArgs A) {
    __shared__ uint32_t stack_rows[kLdsStack * 64];
    uint32_t spill_words[kSpillAlloc];
    LaneStack stk;
    stk.init(stack_rows, spill_words);
    __shared__ unsigned long long cnt_words[kCntWords];
    if (COUNT) count_begin(cnt_words);
    Tracer<MODE, EARLY, COUNT> tr{A.ts, stk, TraceCount{(lds_u64 *)cnt_words, 0u}};
    const ViewParams V = frame_view(A);
    const uint32_t spp = max(1u, V.spp);
    const Env env = find_env(A);
    PixelFeed feed(A);
    bool live = false;
    uint32_t px = 0u, py = 0u, s = 0u, b = 0u, fi = 0u;
    Rng rng{0u};
    AccT<NS> acc;
    FirstHit fh;
    acc_clear(acc, fh);
    F3 ro = fs(0.0f), rd = fs(0.0f), thr = fs(0.0f);
    for (;;) {
        const uint32_t t_it = COUNT ? phase_clock() : 0u;
        if (feed.take(A, live, px, py, fi)) {
            rng = pixel_rng(V, px, py, V.frame + fi);
            acc_clear(acc, fh);
            s = 0u;
            b = 0u;
            camera_ray(V, px, py, rng, ro, rd);
            thr = fs(1.0f);
        }
        if (__ballot(live) == 0ull) break;
        if (live) {
            if (COUNT) count_step(tr.cnt, kPhaseLoop);
            const bool cont = bounce<MODE, EARLY, COUNT>(A, env, tr, b, ro, rd, thr, rng, acc, fh, px, py);
            b++;
            if (!cont || b >= A.max_depth) {
                acc.end_sample();
                s++;
                if (s >= spp) {
                    finalize(A, acc, spp, fh, (size_t)fi * A.frame_stride + (size_t)py * A.ip.width + px);
                    live = false;
                } else {
                    b = 0u;
                    camera_ray(V, px, py, rng, ro, rd);
                    thr = fs(1.0f);
                }
            }
        }
        if (COUNT) count_time(tr.cnt, kPhaseLoop, t_it);
    }
    if (COUNT) {
        count_flush(A.counters, cnt_words);
    }
}

// ─────────────────────────── path queues (wavefront) ───────────────────────────
//
// The same per-pixel computation split at every bounce. A primary pass runs
// sample s of every pixel through its first bounce; paths that go on are
// appended to a queue (wave-aggregated atomics). Then one pass per depth runs
// each queued path through one bounce and compacts the survivors into the
// other queue. Every pass runs full wavefronts of live paths, so the few long
// paths (glass, metal chains) never hold a wavefront of finished pixels. Each
// pixel's RNG stream, its accumulators and every arithmetic step are those of
// gather_pixel, in the same order; only the place where they are kept between
// bounces changes (registers there, HBM here). Samples run one after another,
// as the RNG stream requires.
//
// Each queue is sorted by material as it fills, with two keys: the paths
// whose last bounce was diffuse fill it from the front, those that left a
// specular lobe or a dielectric fill it from the back. The next pass takes
// them in that order, so its wavefronts hold paths of one kind: the long
// glass/metal chains run together instead of one per wavefront of short
// diffuse paths.
//
// Pixel state (SoA planes of n_px float4):
//   0: col.xyz, lum_sum   1: lum2_sum, cur, rng state, 0
//   2: sl[0..3]           3: sl[4..7]           4: first normal.xyz, first depth
// Queue entry (SoA planes): 0: ro.xyz, pixel   1: rd.xyz, rng state   2: thr.xyz, 0
constexpr uint32_t kPixelPlanes = 5;
constexpr uint32_t kQueuePlanes = 3;
constexpr uint32_t kPathCountersPerDepth = 3;

__device__ __forceinline__ void store_running(const GatherArgs &A, uint32_t i, const Acc &acc, uint32_t rng) {
    A.px_state[i] = make_float4(acc.col.x, acc.col.y, acc.col.z, acc.lum_sum);
    A.px_state[A.n_px + i] = make_float4(acc.lum2_sum, acc.cur, __uint_as_float(rng), 0.0f);
}
__device__ __forceinline__ void load_running(const GatherArgs &A, uint32_t i, Acc &acc, uint32_t &rng) {
    const float4 p0 = A.px_state[i], p1 = A.px_state[A.n_px + i];
    acc.col = xyz(p0);
    acc.lum_sum = p0.w;
    acc.lum2_sum = p1.x;
    acc.cur = p1.y;
    rng = __float_as_uint(p1.z);
}
__device__ __forceinline__ void store_samples(const GatherArgs &A, uint32_t i, const Acc &acc, const FirstHit &fh) {
    A.px_state[2u * A.n_px + i] = make_float4(acc.sl[0], acc.sl[1], acc.sl[2], acc.sl[3]);
    A.px_state[3u * A.n_px + i] = make_float4(acc.sl[4], acc.sl[5], acc.sl[6], acc.sl[7]);
    A.px_state[4u * A.n_px + i] = make_float4(fh.normal.x, fh.normal.y, fh.normal.z, fh.depth);
}
__device__ __forceinline__ void load_samples(const GatherArgs &A, uint32_t i, Acc &acc, FirstHit &fh) {
    const float4 a = A.px_state[2u * A.n_px + i], b = A.px_state[3u * A.n_px + i];
    const float4 c = A.px_state[4u * A.n_px + i];
    acc.sl[0] = a.x; acc.sl[1] = a.y; acc.sl[2] = a.z; acc.sl[3] = a.w;
    acc.sl[4] = b.x; acc.sl[5] = b.y; acc.sl[6] = b.z; acc.sl[7] = b.w;
    fh.normal = xyz(c);
    fh.depth = c.w;
}

// Index of this lane's entry when the lanes with `push` append one entry each:
// one atomic per wavefront, ranks from the ballot. All lanes must call it.
__device__ __forceinline__ uint32_t wave_append(bool push, uint32_t *counter) {
    const unsigned long long m = __ballot(push);
    if (m == 0ull) return 0u;
    const uint32_t lane = __lane_id();
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0u;
    if (lane == (uint32_t)leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

__device__ __forceinline__ void queue_put(float4 *q, uint32_t n_px, uint32_t slot, F3 ro, F3 rd, F3 thr,
                                          uint32_t pixel, uint32_t rng) {
    q[slot] = make_float4(ro.x, ro.y, ro.z, __uint_as_float(pixel));
    q[n_px + slot] = make_float4(rd.x, rd.y, rd.z, __uint_as_float(rng));
    q[2u * n_px + slot] = make_float4(thr.x, thr.y, thr.z, 0.0f);
}

// Where a sample's path ended: close the sample; after the last sample write
// the pixel's outputs, else park its state for the next primary pass.
__device__ __forceinline__ void end_path(const GatherArgs &A, uint32_t i, Acc &acc, FirstHit &fh, uint32_t rng,
                                         uint32_t spp) {
    acc.end_sample();
    if (A.sample + 1u >= spp) {
        finalize(A, acc, spp, fh, i);
    } else {
        store_running(A, i, acc, rng);
        store_samples(A, i, acc, fh);
    }
}

template <int MODE, bool EARLY, bool COUNT>
__global__ __launch_bounds__(64) void gather_primary_kernel(GatherArgs A) {
    __shared__ uint32_t stack_rows[kLdsStack * 64];
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t q = nb / 8u;
    const uint32_t tile = (b < 8u * q) ? (b % 8u) * q + b / 8u : b;
    if (tile >= A.n_tiles) return;
    const uint32_t px = (tile % A.tiles_x) * 8u + (threadIdx.x & 7u);
    const uint32_t py = (tile / A.tiles_x) * 8u + (threadIdx.x >> 3);
    uint32_t spill_words[kSpillAlloc];
    LaneStack stk;
    stk.init(stack_rows, spill_words);
    __shared__ unsigned long long cnt_words[kCntWords];
    if (COUNT) count_begin(cnt_words);
    Tracer<MODE, EARLY, COUNT> tr{A.ts, stk, TraceCount{(lds_u64 *)cnt_words, 0u}};
    const ViewParams V = frame_view(A);
    const uint32_t spp = max(1u, V.spp);
    const uint32_t idx = py * A.ip.width + px;
    bool push = false, smooth = false;
    F3 ro, rd, thr;
    Rng rng{0u};
    if (px < A.ip.width && py < A.ip.height) {
        Acc acc;
        FirstHit fh;
        if (A.sample == 0u) {
            rng = pixel_rng(V, px, py);
            acc_clear(acc, fh);
        } else {
            load_running(A, idx, acc, rng.state);
            load_samples(A, idx, acc, fh);
        }
        camera_ray(V, px, py, rng, ro, rd);
        thr = fs(1.0f);
        const Env env = find_env(A);
        const bool cont =
            bounce<MODE, EARLY, COUNT>(A, env, tr, 0u, ro, rd, thr, rng, acc, fh, px, py, &smooth);
        if (cont && A.max_depth > 1u) {
            push = true;
            store_running(A, idx, acc, 0u);
            store_samples(A, idx, acc, fh);
        } else {
            end_path(A, idx, acc, fh, rng.state, spp);
        }
    }
    const uint32_t front = wave_append(push && !smooth, &A.path_count[1]);
    const uint32_t back = wave_append(push && smooth, &A.path_count[2]);
    if (push) queue_put(A.queue[1], A.n_px, smooth ? A.n_px - 1u - back : front, ro, rd, thr, idx, rng.state);
    if (COUNT) {
        count_flush(A.counters, cnt_words);
    }
}

// One bounce of the queued paths at depth b (1 .. max_depth - 1). The paths
// entering depth b are queue[b & 1] (counters path_count[3(b-1) ..]: [0] work
// taken, [1] paths from the front, [2] paths from the back); the paths that go
// on are compacted into queue[(b + 1) & 1] with the same two material keys
// (diffuse from the front, specular or dielectric from the back), counters
// path_count[3b ..]. Wavefronts take 64 paths at a time, front first, so each
// holds paths of one kind, and the finished paths of a bounce leave no idle
// lanes behind. A path's bounces run in the same order with the same state as
// in the one-kernel pass (running sums in the pixel state between bounces,
// ray, throughput and generator in the queue entry), so the bits are the same.
template <int MODE, bool EARLY, bool COUNT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void gather_bounce_kernel(GatherArgs A,
                                                                                                 uint32_t b) {
    __shared__ uint32_t stack_rows[kLdsStack * 64];
    uint32_t spill_words[kSpillAlloc];
    LaneStack stk;
    stk.init(stack_rows, spill_words);
    __shared__ unsigned long long cnt_words[kCntWords];
    if (COUNT) count_begin(cnt_words);
    Tracer<MODE, EARLY, COUNT> tr{A.ts, stk, TraceCount{(lds_u64 *)cnt_words, 0u}};
    uint32_t *cin = A.path_count + 3u * (b - 1u), *cout = A.path_count + 3u * b;
    const uint32_t n_front = __builtin_amdgcn_readfirstlane(cin[1]);
    const uint32_t count = n_front + __builtin_amdgcn_readfirstlane(cin[2]);
    const float4 *qin = A.queue[b & 1u];
    float4 *qout = A.queue[(b + 1u) & 1u];
    const uint32_t spp = max(1u, A.ip.spp);
    const bool last = b + 1u >= A.max_depth;
    const Env env = find_env(A);
    for (;;) {
        uint32_t base = 0u;
        if (threadIdx.x == 0) base = atomicAdd(&cin[0], 64u);
        base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, 64));
        if (base >= count) break;
        const uint32_t w = base + threadIdx.x;
        bool push = false, smooth = false;
        F3 ro, rd, thr;
        Rng rng{0u};
        uint32_t idx = 0u;
        if (w < count) {
            const uint32_t i = w < n_front ? w : A.n_px - 1u - (w - n_front);  // diffuse first, then smooth
            const float4 q0 = qin[i], q1 = qin[A.n_px + i], q2 = qin[2u * A.n_px + i];
            ro = xyz(q0);
            idx = __float_as_uint(q0.w);
            rd = xyz(q1);
            rng.state = __float_as_uint(q1.w);
            thr = xyz(q2);
            Acc acc;
            FirstHit fh{f3(0.0f, 1.0f, 0.0f), 1e30f};  // only depth 0 touches it
            uint32_t unused;
            load_running(A, idx, acc, unused);
            const bool cont = bounce<MODE, EARLY, COUNT>(A, env, tr, b, ro, rd, thr, rng, acc, fh, idx % A.ip.width,
                                                         idx / A.ip.width, &smooth);
            if (cont && !last) {
                push = true;
                store_running(A, idx, acc, 0u);
            } else {
                load_samples(A, idx, acc, fh);
                end_path(A, idx, acc, fh, rng.state, spp);
            }
        }
        const uint32_t front = wave_append(push && !smooth, &cout[1]);
        const uint32_t back = wave_append(push && smooth, &cout[2]);
        if (push) queue_put(qout, A.n_px, smooth ? A.n_px - 1u - back : front, ro, rd, thr, idx, rng.state);
    }
    if (COUNT) {
        count_flush(A.counters, cnt_words);
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
    stk.init(stack_rows, spill_words);
    Tracer<MODE, EARLY, false> tr{B.g.ts, stk, TraceCount{nullptr, 0u}};
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
                    gather_pixel<MODE, EARLY, false>(B.g, V, tr, px, py, bv.pixel_offset + (size_t)py * V.width + px);
            }
        }
    }
}

// The same queue handed out one pixel slot at a time: a slot is (tile of the
// cost-sorted list, pixel inside the tile), and each lane takes the next slot
// when its pixel is done, as gather_persistent_kernel does within one view.
// Each lane reads its own view's record when a pixel or sample starts.
template <int MODE, bool EARLY>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void gather_batch_lanes_kernel(BatchArgs B) {
    __shared__ uint32_t stack_rows[kLdsStack * 64];
    uint32_t spill_words[kSpillAlloc];
    LaneStack stk;
    stk.init(stack_rows, spill_words);
    Tracer<MODE, EARLY, false> tr{B.g.ts, stk, TraceCount{nullptr, 0u}};
    const Env env = find_env(B.g);
    const uint32_t lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    const uint32_t tile_px = B.tile_w * B.tile_h;
    const uint32_t n_slots = B.n_tiles * tile_px;  // < 2^32, checked by the host
    bool more = true, live = false;
    uint32_t vi = 0u, px = 0u, py = 0u, s = 0u, b = 0u;
    Rng rng{0u};
    Acc acc;
    FirstHit fh;
    acc_clear(acc, fh);
    F3 ro = fs(0.0f), rd = fs(0.0f), thr = fs(0.0f);
    for (;;) {
        while (more) {
            const unsigned long long need = __ballot(!live);
            if (need == 0ull) break;
            const uint32_t n_need = (uint32_t)__popcll(need);
            const int leader = __ffsll((long long)need) - 1;
            uint32_t base = 0u;
            if (lane == (uint32_t)leader) base = atomicAdd(B.counter, n_need);
            base = __builtin_amdgcn_readfirstlane(__shfl(base, leader, 64));
            if (!live) {
                const uint32_t k = base + (uint32_t)__popcll(need & below);
                if (k < n_slots) {
                    const uint32_t t = k / tile_px, j = k % tile_px;
                    const uint2 te = B.tiles[t];
                    const BatchView &bv = B.views[te.x];
                    const uint32_t tx = te.y % bv.tiles_x, ty = te.y / bv.tiles_x;
                    px = tx * B.tile_w + j % B.tile_w;
                    py = ty * B.tile_h + j / B.tile_w;
                    if (px < bv.width && py < bv.height) {
                        live = true;
                        vi = te.x;
                        const ViewParams V{bv.cam, bv.width, bv.height, bv.spp, bv.frame};
                        rng = pixel_rng(V, px, py);
                        acc_clear(acc, fh);
                        s = 0u;
                        b = 0u;
                        camera_ray(V, px, py, rng, ro, rd);
                        thr = fs(1.0f);
                    }
                }
            }
            if (base + n_need >= n_slots) more = false;
        }
        if (__ballot(live) == 0ull) break;
        if (live) {
            const bool cont =
                bounce<MODE, EARLY, false>(B.g, env, tr, b, ro, rd, thr, rng, acc, fh, px, py);
            b++;
            if (!cont || b >= B.g.max_depth) {
                acc.end_sample();
                s++;
                const BatchView &bv = B.views[vi];
                const uint32_t spp = max(1u, bv.spp);
                if (s >= spp) {
                    finalize(B.g, acc, spp, fh, bv.pixel_offset + (size_t)py * bv.width + px);
                    live = false;
                } else {
                    const ViewParams V{bv.cam, bv.width, bv.height, bv.spp, bv.frame};
                    b = 0u;
                    camera_ray(V, px, py, rng, ro, rd);
                    thr = fs(1.0f);
                }
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
    stk.init(stack_rows, spill_words);
    TraceCount cnt{nullptr, 0u};
    const rt_ray r = rays[i];
    const F3 ro = ld3(r.origin), rd = ld3(r.dir);
    const Hit h = trace_mode<MODE, false, false>(ts, ro, rd, kInf, stk, cnt);
    out[i].t = h.t;
    out[i].prim_id = h.prim;
}

// First-hit albedo along the pixel-centre ray (zero jitter): a guide image for
// the learned denoiser, not part of the reference's gather.
template <int MODE>
__global__ __launch_bounds__(64) void albedo_kernel(TraceScene ts, const rt_triangle *tris, const rt_material *mats,
                                                    rt_camera_uniform cam, uint32_t w, uint32_t h, float4 *out) {
    __shared__ uint32_t stack_rows[kLdsStack * 64];
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= w * h) return;
    uint32_t spill_words[kSpillAlloc];
    LaneStack stk;
    stk.init(stack_rows, spill_words);
    TraceCount cnt{nullptr, 0u};
    const uint32_t px = i % w, py = i / w;
    const float ux = (float)px * cam.inv_half_res[0] - 1.0f;
    const float uy = 1.0f - (float)py * cam.inv_half_res[1];
    const F3 ro = ld3(cam.position);
    const F3 rd = normalize(ld3(cam.forward) + ld3(cam.right) * ux + ld3(cam.up) * uy);
    const Hit hit = trace_mode<MODE, false, false>(ts, ro, rd, kInf, stk, cnt);
    float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (hit.t < kInf) {
        const rt_material &m = mats[tris[hit.prim].material_id];
        o = make_float4(m.albedo[0], m.albedo[1], m.albedo[2], 1.0f);
    }
    out[i] = o;
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

namespace rtd {

// Slot k of the watertight traversal: the vertices of triangle prims[k] as
// uploaded (no edge subtraction, so shared vertices stay identical).
__global__ __launch_bounds__(256) void wt_slots_kernel(const rt_triangle *__restrict__ tris,
                                                       const uint32_t *__restrict__ prims, uint32_t n,
                                                       float4 *__restrict__ out) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= n) return;
    const uint32_t p = prims[k];
    const rt_triangle &t = tris[p];
    out[3 * k] = make_float4(t.v0[0], t.v0[1], t.v0[2], __uint_as_float(p));
    out[3 * k + 1] = make_float4(t.v1[0], t.v1[1], t.v1[2], 0.0f);
    out[3 * k + 2] = make_float4(t.v2[0], t.v2[1], t.v2[2], 0.0f);
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
    A.gather_rec = reinterpret_cast<const float4 *>(g.gather_rec);
    A.frame = reinterpret_cast<float4 *>(g.frame);
    A.moments = reinterpret_cast<float4 *>(g.moments);
    A.gbuffer = reinterpret_cast<float4 *>(g.gbuffer);
    A.max_depth = g.max_depth;
    A.tiles_x = (g.ip.width + 7u) / 8u;
    const uint32_t tiles_y = (g.ip.height + 7u) / 8u;
    A.n_tiles = A.tiles_x * tiles_y;
    A.counters = g.counters;
    A.px_state = reinterpret_cast<float4 *>(g.px_state);
    A.queue[0] = reinterpret_cast<float4 *>(g.queue[0]);
    A.queue[1] = reinterpret_cast<float4 *>(g.queue[1]);
    A.path_count = g.path_count;
    A.n_px = g.ip.width * g.ip.height;
    A.sample = 0u;
    for (uint32_t k = 0; k < RT_TEX_ROLES; k++) A.tex[k] = g.tex[k];
    A.regen = g.regen;
    A.refill_min = std::max(g.refill_min, 1u);
    A.fstate = g.fstate;
    A.n_frames = std::max(g.n_frames, 1u);
    A.frame_stride = g.frame_stride;
    return A;
}

// Workgroups of the persistent gather that are resident at once on the
// current device (one 64-lane wavefront each).
// The count is computed once per instance (a function-local static:
// initialised once even when renderers on several threads launch first).
template <class K>
static uint32_t resident_blocks(K kernel) {
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, 0) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0u;
    return (uint32_t)std::max(per_cu, 1) * (uint32_t)std::max(cus, 1);
}

template <int M, bool E, bool C, int NS>
static uint32_t persistent_blocks() {
    static const uint32_t blocks = resident_blocks(rtd::gather_persistent_kernel<M, E, C, NS>);
    return blocks;
}

template <int M, bool E, bool C, int NS>
static hipError_t launch_persistent_ns(const rtd::GatherArgs &A, uint32_t max_blocks, hipStream_t s) {
    uint32_t nb = persistent_blocks<M, E, C, NS>();
    if (nb == 0u) return hipErrorInvalidValue;
    if (max_blocks) nb = std::min(nb, max_blocks);
    // more workgroups than pixel slots would only start and leave
    nb = std::min(nb, std::max(A.n_tiles, 1u));
    const hipError_t e = launch_fill_u32(A.regen, 0u, rtd::kBands, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((rtd::gather_persistent_kernel<M, E, C, NS>), dim3(nb), dim3(64), 0, s, A);
    return hipSuccess;
}

// 1-spp frames of the default traversal use the one-sample accumulator
// (fewer live registers on the whole path); everything else the 8-sample one.
template <int M, bool E, bool C>
static hipError_t launch_persistent(const rtd::GatherArgs &A, uint32_t max_blocks, hipStream_t s) {
    if constexpr (M == 3) {
        if (A.ip.spp <= 1u) return launch_persistent_ns<M, E, C, 1>(A, max_blocks, s);
    }
    return launch_persistent_ns<M, E, C, 8>(A, max_blocks, s);
}

size_t gather_wavefront_bytes(uint32_t n_px, uint32_t max_depth) {
    return (size_t)n_px * 16u * (rtd::kPixelPlanes + 2u * rtd::kQueuePlanes) +
           4u * rtd::kPathCountersPerDepth * ((size_t)std::max(max_depth, 1u) + 1u);
}

template <int M, bool E, bool C>
static void launch_wavefront_passes(rtd::GatherArgs A, hipStream_t s) {
    const uint32_t spp = A.ip.spp > 0u ? A.ip.spp : 1u;
    // four one-wavefront workgroups per SIMD of 256 CUs, as many as are resident
    const uint32_t bounce_wgs = std::min<uint32_t>((A.n_px + 63u) / 64u, 4096u);
    const uint32_t n_counters = rtd::kPathCountersPerDepth * (std::max(A.max_depth, 1u) + 1u);
    for (uint32_t k = 0; k < spp; k++) {
        A.sample = k;
        (void)launch_fill_u32(A.path_count, 0u, n_counters, s);
        hipLaunchKernelGGL((rtd::gather_primary_kernel<M, E, C>), dim3(A.n_tiles), dim3(64), 0, s, A);
        // one launch per depth; each reads its queue's size from device memory
        for (uint32_t b = 1u; b < A.max_depth; b++)
            hipLaunchKernelGGL((rtd::gather_bounce_kernel<M, E, C>), dim3(bounce_wgs), dim3(64), 0, s, A, b);
    }
}

hipError_t launch_gather(const GatherLaunch &g, hipStream_t s) {
    const rtd::GatherArgs A = gather_args(g);
    const dim3 grid(A.n_tiles), block(64);
#define RT_GATHER(M, E, C)                                                                    \
    do {                                                                                      \
        if (g.px_state) launch_wavefront_passes<M, E, C>(A, s);                              \
        else if (g.regen) {                                                                   \
            const hipError_t e = launch_persistent<M, E, C>(A, g.regen_blocks, s);            \
            if (e != hipSuccess) return e;                                                    \
        } else hipLaunchKernelGGL((rtd::gather_kernel<M, E, C>), grid, block, 0, s, A);      \
    } while (0)
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
    else if (g.mode == 4) { RT_GATHER_M(4) }
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
#define RT_BATCH(M, E)                                                                          \
    do {                                                                                        \
        if (b.lanes) hipLaunchKernelGGL((rtd::gather_batch_lanes_kernel<M, E>), grid, block, 0, s, B); \
        else hipLaunchKernelGGL((rtd::gather_batch_kernel<M, E>), grid, block, 0, s, B);        \
    } while (0)
    if (g.mode == 0) { if (g.early_shadow) RT_BATCH(0, true); else RT_BATCH(0, false); }
    else if (g.mode == 1) { if (g.early_shadow) RT_BATCH(1, true); else RT_BATCH(1, false); }
    else if (g.mode == 2) { if (g.early_shadow) RT_BATCH(2, true); else RT_BATCH(2, false); }
    else if (g.mode == 4) { if (g.early_shadow) RT_BATCH(4, true); else RT_BATCH(4, false); }
    else { if (g.early_shadow) RT_BATCH(3, true); else RT_BATCH(3, false); }
#undef RT_BATCH
    return hipGetLastError();
}

hipError_t launch_albedo(const TraceInputs &t, uint32_t mode, const rt_triangle *tris, const rt_material *mats,
                         const rt_camera_uniform &cam, uint32_t w, uint32_t h, void *out, hipStream_t s) {
    const uint32_t n = w * h;
    if (n == 0) return hipSuccess;
    const rtd::TraceScene ts = rtd::make_trace_scene(t);
    float4 *o = reinterpret_cast<float4 *>(out);
    const dim3 grid((n + 63u) / 64u), block(64);
    if (mode == 0) hipLaunchKernelGGL((rtd::albedo_kernel<0>), grid, block, 0, s, ts, tris, mats, cam, w, h, o);
    else if (mode == 3) hipLaunchKernelGGL((rtd::albedo_kernel<3>), grid, block, 0, s, ts, tris, mats, cam, w, h, o);
    else if (mode == 4) hipLaunchKernelGGL((rtd::albedo_kernel<4>), grid, block, 0, s, ts, tris, mats, cam, w, h, o);
    else hipLaunchKernelGGL((rtd::albedo_kernel<1>), grid, block, 0, s, ts, tris, mats, cam, w, h, o);
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
    else if (mode == 4) hipLaunchKernelGGL((rtd::trace_rays_kernel<4>), grid, block, 0, s, ts, rays, n, out);
    else hipLaunchKernelGGL((rtd::trace_rays_kernel<3>), grid, block, 0, s, ts, rays, n, out);
    return hipGetLastError();
}

hipError_t launch_wt_slots(const rt_triangle *tris, const uint32_t *prims, uint32_t n_slots, void *out,
                           hipStream_t s) {
    if (n_slots == 0) return hipSuccess;
    hipLaunchKernelGGL(rtd::wt_slots_kernel, dim3((n_slots + 255u) / 256u), dim3(256), 0, s, tris, prims, n_slots,
                       reinterpret_cast<float4 *>(out));
    return hipGetLastError();
}

}  // namespace rtk
