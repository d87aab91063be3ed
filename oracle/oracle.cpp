// oracle.cpp — CPU restatement of rustracer's WGSL kernels.
//
// TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker; never by the product.
//
// Restates, line by line (reference paths relative to the repository root):
//   rustracer-renderer/shaders/common.wgsl      (all)
//   rustracer-renderer/shaders/gather.wgsl      (all; textures sampled as
//                                                 include/rt_texture.h fixes it)
//   rustracer-renderer/shaders/photon_trace.wgsl
//   rustracer-renderer/shaders/photon_count.wgsl + CPU scan lib.rs:987-1007
//   rustracer-renderer/shaders/photon_scatter.wgsl (made stable: ascending slot)
//   rustracer-renderer/shaders/denoise.wgsl, composite.wgsl, tonemap.wgsl
//   rustracer-app/examples/common/mod.rs:37-47 (CPU tonemap of save_png)
// Beyond the reference: the watertight ray/triangle test of the opt-in
// watertight traversal (Woop, Benthin, Wald, JCGT 2013), as a brute-force
// closest hit, so the GPU mode can be checked bit for bit.
// Conscious divergences (identical in the HIP backend):
//   * an out-of-range light index (rng_f32 == 1.0, gather.wgsl:155,
//     photon_trace.wgsl:113) is clamped to n-1, i.e. wgpu's "Restrict" policy;
//   * the loop bound `b < 6u` (gather.wgsl:299, photon_trace.wgsl:178) is the
//     parameter max_depth (6 reproduces the reference);
//   * within a grid cell photons are ordered by ascending slot index, where the
//     reference keeps atomic arrival order (photon_scatter.wgsl:35).
// Parity pins: see tests/test_oracle_*.py (closed forms, brute force, libm);
// the normal-map tangent frame (common.wgsl tangent_frame, gather.wgsl:330-336)
// by closed forms of its debug views on quads with known uv layouts.

#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../include/rt_wgsl_math.h"

namespace {

constexpr float PI = 3.141592653589793f;  // common.wgsl:8
constexpr float INF = 1e38f;              // common.wgsl:9
constexpr float EPS = 1e-6f;              // common.wgsl:10

struct V3 {
    float x, y, z;
};
struct V2 {
    float x, y;
};

inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 splat(float s) { return V3{s, s, s}; }
inline V3 L3(const float *p) { return V3{p[0], p[1], p[2]}; }
inline V2 L2(const float *p) { return V2{p[0], p[1]}; }
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
inline V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
inline V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
inline V2 operator+(V2 a, V2 b) { return V2{a.x + b.x, a.y + b.y}; }
inline V2 operator-(V2 a, V2 b) { return V2{a.x - b.x, a.y - b.y}; }
inline V2 operator*(V2 a, float s) { return V2{a.x * s, a.y * s}; }
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float length(V3 v) { return wgsl_sqrt(dot(v, v)); }
inline V3 normalize(V3 v) { return v * (1.0f / wgsl_sqrt(dot(v, v))); }
inline V3 vmin(V3 a, V3 b) { return V3{wmin(a.x, b.x), wmin(a.y, b.y), wmin(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return V3{wmax(a.x, b.x), wmax(a.y, b.y), wmax(a.z, b.z)}; }
// mix(a, b, t) = a * (1 - t) + b * t (WGSL builtin definition)
inline V3 mix(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
inline V3 mix(V3 a, V3 b, V3 t) { return a * (splat(1.0f) - t) + b * t; }
inline float vmax3(V3 v) { return wmax(wmax(v.x, v.y), v.z); }
inline V3 vpow(V3 v, float e) { return V3{wgsl_pow(v.x, e), wgsl_pow(v.y, e), wgsl_pow(v.z, e)}; }

// ── common.wgsl:12-25 ──
inline V3 reflect(V3 v, V3 n) { return v - (2.0f * dot(v, n)) * n; }
inline V3 refract(V3 v, V3 n, float eta) {
    const float cos_theta = wmin(dot(-v, n), 1.0f);
    const V3 r_out_perp = eta * (v + cos_theta * n);
    const V3 r_out_parallel = (-wgsl_sqrt(std::fabs(1.0f - dot(r_out_perp, r_out_perp)))) * n;
    return r_out_perp + r_out_parallel;
}
inline V3 fresnel_schlick(float cos_theta, V3 f0) {
    return f0 + (splat(1.0f) - f0) * wgsl_pow(wclamp(1.0f - cos_theta, 0.0f, 1.0f), 5.0f);
}

// ── PCG, common.wgsl:29-56 ──
struct Rng {
    uint32_t state, inc;
};
inline Rng rng_init(uint32_t seed) {
    Rng r;
    r.state = seed;
    r.inc = 1u;
    r.state = r.state * 747796405u + 2891336453u;
    r.state = ((r.state >> ((r.state >> 28u) + 4u)) ^ r.state) * 277803737u;
    return r;
}
inline uint32_t rng_next(Rng *r) {
    const uint32_t old_state = r->state;
    r->state = old_state * 747796405u + r->inc;
    const uint32_t word = ((old_state >> ((old_state >> 28u) + 4u)) ^ old_state) * 277803737u;
    return (word >> 22u) ^ word;
}
inline float rng_f32(Rng *r) { return (float)rng_next(r) / 4294967296.0f; }

// ── common.wgsl:60-74 ──
inline V3 sample_cosine_hemisphere(Rng *r) {
    const float u1 = rng_f32(r);
    const float u2 = rng_f32(r);
    const float phi = (2.0f * PI) * u1;
    const float cos_theta = wgsl_sqrt(u2);
    const float sin_theta = wgsl_sqrt(1.0f - u2);
    return v3(wgsl_cos(phi) * sin_theta, cos_theta, wgsl_sin(phi) * sin_theta);
}
inline V3 to_world(V3 local_dir, V3 normal) {
    const V3 up = (std::fabs(normal.y) < 0.999f) ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    const V3 tangent = normalize(cross(up, normal));
    const V3 bitangent = cross(normal, tangent);
    return tangent * local_dir.x + normal * local_dir.y + bitangent * local_dir.z;
}

// ── Möller–Trumbore, common.wgsl:78-100 ──
inline float intersect_triangle(V3 ro, V3 rd, V3 v0, V3 v1, V3 v2) {
    const V3 edge1 = v1 - v0;
    const V3 edge2 = v2 - v0;
    const V3 h = cross(rd, edge2);
    const float a = dot(edge1, h);
    if (std::fabs(a) < 1e-7f) return INF;
    const float f = 1.0f / a;
    const V3 s = ro - v0;
    const float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return INF;
    const V3 q = cross(s, edge1);
    const float v = f * dot(rd, q);
    if (v < 0.0f || u + v > 1.0f) return INF;
    const float t = f * dot(edge2, q);
    if (t > EPS) return t;
    return INF;
}

// ── Watertight ray/triangle test (opt-in mode; not in the reference) ──
// Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013,
// section 3: permute so the largest |direction| axis is z (swapping x and y
// when that component is negative, which keeps the winding), shear the
// vertices into ray space, take the three 2-D edge functions, and recompute
// them in double when one is exactly zero. Float32, no contraction; two-sided,
// with the reference's t > 1e-6.
inline float axis(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
inline float intersect_watertight(V3 ro, V3 rd, V3 v0, V3 v1, V3 v2) {
    const float mx = std::fabs(rd.x), my = std::fabs(rd.y), mz = std::fabs(rd.z);
    const int kz = (mx >= my && mx >= mz) ? 0 : (my >= mz ? 1 : 2);
    int kx = (kz + 1) % 3, ky = (kz + 2) % 3;
    const float dz = axis(rd, kz);
    if (dz < 0.0f) std::swap(kx, ky);
    const float shx = axis(rd, kx) / dz, shy = axis(rd, ky) / dz, shz = 1.0f / dz;
    const V3 p[3] = {v0 - ro, v1 - ro, v2 - ro};
    float x[3], y[3], z[3];
    for (int i = 0; i < 3; i++) {
        z[i] = axis(p[i], kz);
        x[i] = axis(p[i], kx) - shx * z[i];
        y[i] = axis(p[i], ky) - shy * z[i];
    }
    // edge function opposite each vertex: (1,2) for vertex 0, (2,0), (0,1)
    float e[3];
    for (int i = 0; i < 3; i++) {
        const int a = (i + 2) % 3, b = (i + 1) % 3;
        e[i] = x[a] * y[b] - y[a] * x[b];
    }
    if (e[0] == 0.0f || e[1] == 0.0f || e[2] == 0.0f) {
        for (int i = 0; i < 3; i++) {
            const int a = (i + 2) % 3, b = (i + 1) % 3;
            e[i] = (float)((double)x[a] * (double)y[b] - (double)y[a] * (double)x[b]);
        }
    }
    const bool any_neg = e[0] < 0.0f || e[1] < 0.0f || e[2] < 0.0f;
    const bool any_pos = e[0] > 0.0f || e[1] > 0.0f || e[2] > 0.0f;
    if (any_neg && any_pos) return INF;
    const float det = (e[0] + e[1]) + e[2];
    if (det == 0.0f) return INF;
    const float tsc = (e[0] * (shz * z[0]) + e[1] * (shz * z[1])) + e[2] * (shz * z[2]);
    const float t = tsc / det;
    return t > EPS ? t : INF;
}

// ── common.wgsl:171-251 ──
struct Bary {
    V2 uv;
    float u, v;
};
inline Bary barycentric(const rt_triangle &tri, V3 ro, V3 rd, float t) {
    const V3 v0 = L3(tri.v0);
    const V3 e1 = L3(tri.v1) - v0;
    const V3 e2 = L3(tri.v2) - v0;
    const V3 p = (ro + rd * t) - v0;
    const float d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2);
    const float d20 = dot(p, e1), d21 = dot(p, e2);
    const float denom = d00 * d11 - d01 * d01;
    float u = 0.0f, v = 0.0f;
    if (denom != 0.0f) {
        u = (d11 * d20 - d01 * d21) / denom;
        v = (d00 * d21 - d01 * d20) / denom;
    }
    const float w = 1.0f - u - v;
    return Bary{L2(tri.uv0) * w + L2(tri.uv1) * u + L2(tri.uv2) * v, u, v};
}
inline V3 smooth_normal(const rt_triangle &tri, const Bary &b, V3 geometric) {
    const float w = 1.0f - b.u - b.v;
    const V3 n = L3(tri.n0) * w + L3(tri.n1) * b.u + L3(tri.n2) * b.v;
    if (dot(n, n) < 1e-8f) return geometric;
    return normalize(n);
}
struct TanFrame {
    V3 t, b;
};
inline TanFrame tangent_frame(const rt_triangle &tri, V3 n) {
    const V3 v0 = L3(tri.v0);
    const V3 e1 = L3(tri.v1) - v0;
    const V3 e2 = L3(tri.v2) - v0;
    const V2 d1 = L2(tri.uv1) - L2(tri.uv0);
    const V2 d2 = L2(tri.uv2) - L2(tri.uv0);
    const float det = d1.x * d2.y - d2.x * d1.y;
    const float f = (std::fabs(det) < 1e-12f) ? 0.0f : 1.0f / det;
    const V3 t = (e1 * d2.y - e2 * d1.y) * f;
    const V3 b = (e2 * d1.x - e1 * d2.x) * f;
    V3 tt = t - n * dot(n, t);
    if (dot(tt, tt) < 1e-10f) {
        tt = cross(n, v3(1.0f, 0.0f, 0.0f));
        if (dot(tt, tt) < 1e-6f) tt = cross(n, v3(0.0f, 1.0f, 0.0f));
    }
    tt = normalize(tt);
    const V3 bb = cross(n, tt) * ((dot(cross(n, tt), b) > 0.0f) ? 1.0f : -1.0f);
    return TanFrame{tt, bb};
}
inline V3 get_triangle_normal(const rt_triangle &tri) {
    const V3 v0 = L3(tri.v0);
    return normalize(cross(L3(tri.v1) - v0, L3(tri.v2) - v0));
}

struct Hit {
    float t;
    uint32_t prim_id;
};

// Everything one invocation reads, as the bindings of gather.wgsl:10-28.
struct Ctx {
    const or_scene *s;
    const rt_image_params *ip;
    const rt_camera_uniform *cam;
    const or_photon_map *pm;
    uint32_t max_depth;
    or_stats st;
};

inline uint32_t clamp_index(uint32_t i, uint32_t n, uint64_t *overflow) {
    if (i >= n) {
        if (overflow) ++*overflow;
        return n - 1u;
    }
    return i;
}

// ── trace_bvh, gather.wgsl:66-88 (twin: photon_trace.wgsl:31-81) ──
Hit trace_bvh(const or_scene *s, V3 ro, V3 rd, or_stats *st) {
    Hit hit;
    hit.t = INF;
    hit.prim_id = 0u;
    const V3 inv = splat(1.0f) / rd;
    uint32_t idx = 0u;
    uint32_t stk[64];
    uint32_t sp = 0u;
    if (st) st->ray_queries++;
    if (s->n_nodes == 0) return hit;  // empty scene: the reference binds one zeroed node
    for (;;) {
        const rt_bvh_node &n = s->nodes[idx];
        if (st) st->nodes_visited++;
        const V3 t0 = (L3(n.bbox_min) - ro) * inv;
        const V3 t1 = (L3(n.bbox_max) - ro) * inv;
        const V3 tm = vmin(t0, t1);
        const V3 tx = vmax(t0, t1);
        const float te = wmax(wmax(tm.x, tm.y), tm.z);
        const float ta = wmin(wmin(tx.x, tx.y), tx.z);
        if (te <= ta && ta > 0.0f && te < hit.t) {
            if (n.prim_count > 0u) {
                for (uint32_t i = 0u; i < n.prim_count; i++) {
                    const uint32_t pi = s->prims[n.left_first + i];
                    const rt_triangle &tr = s->tris[pi];
                    if (st) st->tris_tested++;
                    const float t = intersect_triangle(ro, rd, L3(tr.v0), L3(tr.v1), L3(tr.v2));
                    if (t < hit.t) {
                        hit.t = t;
                        hit.prim_id = pi;
                    }
                }
                if (sp == 0u) break;
                sp -= 1u;
                idx = stk[sp];
            } else {
                if (sp < 63u) {
                    stk[sp] = n.right_child;
                    sp += 1u;
                }
                idx = n.left_first;
            }
        } else {
            if (sp == 0u) break;
            sp -= 1u;
            idx = stk[sp];
        }
    }
    return hit;
}

// textureSampleLevel(role, sampler, uv, i32(layer), 0.0) (include/rt_texture.h)
struct V4 {
    float x, y, z, w;
    V3 rgb() const { return v3(x, y, z); }
};
inline V4 tex(const or_scene *s, uint32_t role, V2 uv, uint32_t layer) {
    float o[4];
    rt_tex_sample(&s->tex[role], uv.x, uv.y, layer, o);
    return V4{o[0], o[1], o[2], o[3]};
}

// gather.wgsl:91-103: the closest hit decides; a transparent texel of an
// alpha-cutout material lets the shadow ray through.
bool shadowed(Ctx &c, V3 origin, V3 dir, float max_dist) {
    const Hit sh = trace_bvh(c.s, origin, dir, &c.st);
    if (sh.t >= max_dist - 0.001f) return false;
    const rt_triangle &bt = c.s->tris[sh.prim_id];
    const rt_material &bm = c.s->mats[bt.material_id];
    if ((bm.flags & RT_MAT_FLAG_ALPHA_CUTOUT) != 0u && bm.tex_albedo != RT_TEX_NONE) {
        const V2 uv = barycentric(bt, origin, dir, sh.t).uv;
        const float alpha = tex(c.s, RT_TEX_ROLE_ALBEDO, uv, bm.tex_albedo).w;
        if (alpha < 0.5f) return false;
    }
    return true;
}

// gather.wgsl:119-141
struct MatParams {
    V3 albedo;
    float rough, metal;
};
inline MatParams eval_material(const or_scene *s, const rt_material &mat, V2 uv) {
    MatParams p{v3(mat.albedo[0], mat.albedo[1], mat.albedo[2]), mat.roughness, mat.metallic};
    if (mat.tex_albedo != RT_TEX_NONE) p.albedo = p.albedo * tex(s, RT_TEX_ROLE_ALBEDO, uv, mat.tex_albedo).rgb();
    if (mat.tex_mr != RT_TEX_NONE) {
        const V4 mr = tex(s, RT_TEX_ROLE_MR, uv, mat.tex_mr);
        p.rough = wclamp(mat.roughness * mr.y, 0.0f, 1.0f);
        p.metal = wclamp(mat.metallic * mr.z, 0.0f, 1.0f);
    }
    return p;
}
inline V3 emissive_at(const or_scene *s, const rt_material &mat, V2 uv) {
    if (mat.tex_emissive != RT_TEX_NONE) return L3(mat.emissive) * tex(s, RT_TEX_ROLE_EMISSIVE, uv, mat.tex_emissive).rgb();
    return L3(mat.emissive);
}

// gather.wgsl:107-111
inline V3 sky_color(V3 dir, V3 zenith, V3 horizon, float intensity) {
    const float t = wclamp(dir.y * 0.5f + 0.5f, 0.0f, 1.0f);
    return mix(horizon, zenith, wgsl_pow(t, 0.5f)) * intensity;
}

struct LightSample {
    V3 e, wi;
};

// gather.wgsl:152-206
LightSample direct_light(Ctx &c, V3 hp, V3 sn, Rng *rng) {
    const uint32_t n = c.s->n_lights;
    if (n == 0u) return LightSample{splat(0.0f), sn};
    const uint32_t li = clamp_index(wgsl_u32(rng_f32(rng) * (float)n), n, &c.st.light_index_overflow);
    const rt_light &l = c.s->lights[li];
    const V3 ldata = L3(l.data);
    const V3 lcolor = L3(l.color);
    if (l.kind == 0u) {  // point
        const V3 to = ldata - hp;
        const float r2 = dot(to, to);
        const float r = wgsl_sqrt(r2);
        const V3 ld = to / r;
        if (dot(sn, ld) <= 0.0f) return LightSample{splat(0.0f), ld};
        if (shadowed(c, hp + sn * 0.01f, ld, r - 0.001f)) return LightSample{splat(0.0f), ld};
        const V3 ple = L3(c.ip->point_light_emission);
        return LightSample{lcolor * l.intensity * ple * wmax(dot(sn, ld), 0.0f) / (r2 + 0.01f) *
                               (float)n,
                           ld};
    } else if (l.kind == 1u) {  // directional
        const V3 ld = normalize(ldata);
        const float ndl = dot(sn, ld);
        if (ndl <= 0.0f) return LightSample{splat(0.0f), ld};
        if (shadowed(c, hp + sn * 0.01f, ld, 1e30f)) return LightSample{splat(0.0f), ld};
        return LightSample{lcolor * l.intensity * ndl * (float)n, ld};
    } else if (l.kind == 2u) {  // area (emissive triangle)
        const uint32_t ti = clamp_index(wgsl_u32(l.data[0]), c.s->n_tris, nullptr);
        const rt_triangle &tri = c.s->tris[ti];
        const float u1 = rng_f32(rng);
        const float u2 = rng_f32(rng);
        const float su = wgsl_sqrt(u1);
        const float u = 1.0f - su;
        const float v = u2 * su;
        const float w = 1.0f - u - v;
        const V3 lp = L3(tri.v0) * w + L3(tri.v1) * u + L3(tri.v2) * v;
        const rt_material &lmat = c.s->mats[tri.material_id];
        const V2 luv = L2(tri.uv0) * w + L2(tri.uv1) * u + L2(tri.uv2) * v;
        V3 le = lcolor;
        if (lmat.tex_emissive != RT_TEX_NONE) le = le * tex(c.s, RT_TEX_ROLE_EMISSIVE, luv, lmat.tex_emissive).rgb();
        le = le * L3(c.ip->light_emission);
        const V3 to = lp - hp;
        const float r2 = dot(to, to);
        const float r = wgsl_sqrt(r2);
        const V3 ld = to / r;
        const float ndl = dot(sn, ld);
        if (ndl <= 0.0f) return LightSample{splat(0.0f), ld};
        const V3 ln = get_triangle_normal(tri);
        const float ndl_l = dot(ln, -ld);
        if (ndl_l <= 0.0f) return LightSample{splat(0.0f), ld};
        if (shadowed(c, hp + sn * 0.01f, ld, r - 0.001f)) return LightSample{splat(0.0f), ld};
        const float area = 0.5f * length(cross(L3(tri.v1) - L3(tri.v0), L3(tri.v2) - L3(tri.v0)));
        return LightSample{le * l.intensity * ndl * ndl_l * area * (float)n / (r2 + 0.01f), ld};
    }
    return LightSample{splat(0.0f), sn};
}

// gather.wgsl:210-227
inline float smith_g1(float nd, float a) {
    const float k = a * 0.5f;
    return nd / (nd * (1.0f - k) + k);
}
inline V3 spec_ggx(V3 wo, V3 n, V3 wi, float a, V3 f0) {
    const float ndl = dot(n, wi);
    const float ndv = dot(n, wo);
    if (ndl <= 0.0f || ndv <= 0.0f) return splat(0.0f);
    const V3 h = normalize(wo + wi);
    const float ndh = wmax(dot(n, h), 0.0f);
    const float a2 = a * a;
    const float dd = ndh * ndh * (a2 - 1.0f) + 1.0f;
    const float D = a2 / (PI * dd * dd);
    const V3 F = fresnel_schlick(wmax(dot(wo, h), 0.0f), f0);
    return F * (D * smith_g1(ndl, a) * smith_g1(ndv, a) / (4.0f * ndl * ndv));
}

// gather.wgsl:230-256. Columns of basis_z are (t, cross(n, t), n).
inline V3 sample_spec_dir(V3 rd, V3 n, float a, Rng *rng) {
    const V3 up = (std::fabs(n.y) < 0.999f) ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    const V3 bt = normalize(cross(up, n));
    const V3 bb = cross(n, bt);
    const V3 mrd = -rd;
    const V3 v = v3(dot(bt, mrd), dot(bb, mrd), dot(n, mrd));  // transpose(B) * (-rd)
    const V3 vh = normalize(v3(a * v.x, a * v.y, wmax(v.z, 1e-4f)));
    const float l2 = vh.x * vh.x + vh.y * vh.y;
    const V3 t1 = (l2 > 0.0f) ? v3(-vh.y, vh.x, 0.0f) * wgsl_inverse_sqrt(l2) : v3(1.0f, 0.0f, 0.0f);
    const V3 t2 = cross(vh, t1);
    const float r = wgsl_sqrt(rng_f32(rng));
    const float phi = (2.0f * PI) * rng_f32(rng);
    const float x = r * wgsl_cos(phi);
    const float y = r * wgsl_sin(phi);
    const float s = 0.5f * (1.0f + vh.z);
    const float yy = (1.0f - s) * wgsl_sqrt(wmax(0.0f, 1.0f - x * x)) + s * y;
    const V3 hh = x * t1 + yy * t2 + wgsl_sqrt(wmax(0.0f, 1.0f - x * x - yy * yy)) * vh;
    const V3 hn = normalize(hh);
    const V3 m = bt * hn.x + bb * hn.y + n * hn.z;  // B * normalize(hh)
    const V3 out = reflect(rd, m);
    // select(out, reflect(rd, n), dot(out, n) > 0): the mirror is taken when
    // the sampled direction is ABOVE the surface (reference behaviour).
    return (dot(out, n) > 0.0f) ? reflect(rd, n) : out;
}

inline float luma_of(V3 c) { return dot(c, v3(0.2126f, 0.7152f, 0.0722f)); }

struct PixelAcc {
    V3 col;
    float lum_sum, lum2_sum;
    float sample_lum[8];
};
// gather.wgsl:36-42
inline void add_radiance(uint32_t s, PixelAcc &p, V3 c) {
    p.col = p.col + c;
    const float l = luma_of(c);
    p.lum_sum += l;
    p.lum2_sum += l * l;
    if (s < 8u) p.sample_lum[s] += l;
}

// gather.wgsl:55-64
struct I3 {
    int32_t x, y, z;
};
inline I3 photon_cell_i(const rt_grid_params *gp, V3 pos) {
    return I3{wgsl_i32((pos.x - gp->room_min_x) / gp->cell_size),
              wgsl_i32((pos.y - gp->room_min_y) / gp->cell_size),
              wgsl_i32((pos.z - gp->room_min_z) / gp->cell_size)};
}

inline uint32_t cell_index(const rt_grid_params *gp, I3 c) {
    return (uint32_t)c.x + (uint32_t)c.y * gp->grid_x + (uint32_t)c.z * gp->grid_x * gp->grid_y;
}
inline int32_t iclamp(int32_t v, int32_t lo, int32_t hi) { return std::min(std::max(v, lo), hi); }
inline int32_t iadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

// gather.wgsl:473-514: the photons of the 27 cells around hp (indices clamped
// into the grid, so border cells repeat), in cell order and stored order; at
// most 384 candidates are read and the search stops at the 64th photon within
// the radius. Returns sum(power * max(n . -incident, 0)) over those photons and
// their count; *reads counts the candidates read.
static V3 photon_estimate(const or_photon_map *pm, V3 hp, V3 shn, float radius, uint32_t *hits_out,
                          uint64_t *reads) {
    const rt_grid_params *gp = pm->gp;
    const float r2 = radius * radius;
    V3 ph_sum = splat(0.0f);
    uint32_t ph_hits = 0u;
    uint32_t budget = 384u;
    bool done = false;
    const I3 base = photon_cell_i(gp, hp);
    const int32_t gxmax = (int32_t)gp->grid_x - 1;
    const int32_t gymax = (int32_t)gp->grid_y - 1;
    const int32_t gzmax = (int32_t)gp->grid_z - 1;
    for (int32_t dz = -1; dz <= 1 && !done; dz++) {
        for (int32_t dy = -1; dy <= 1 && !done; dy++) {
            for (int32_t dx = -1; dx <= 1 && !done; dx++) {
                const I3 cc{iclamp(iadd(base.x, dx), 0, gxmax), iclamp(iadd(base.y, dy), 0, gymax),
                            iclamp(iadd(base.z, dz), 0, gzmax)};
                const uint32_t cell = cell_index(gp, cc);
                const uint32_t start = pm->grid_meta[2 * cell];
                const uint32_t count = pm->grid_meta[2 * cell + 1];
                for (uint32_t i = 0u; i < count && !done; i++) {
                    if (budget == 0u) {
                        done = true;
                        break;
                    }
                    budget -= 1u;
                    const rt_photon &ph = pm->sorted[start + i];
                    (*reads)++;
                    const V3 pp = L3(ph.position);
                    const float d2 = dot(pp - hp, pp - hp);
                    if (d2 < r2) {
                        ph_sum = ph_sum + L3(ph.power) * wmax(dot(shn, -L3(ph.incident)), 0.0f);
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
    *hits_out = ph_hits;
    return ph_sum;
}


// gather.wgsl:258-594, one invocation.
void gather_pixel(Ctx &c, uint32_t px, uint32_t py, float *frame, float *moments, float *gbuffer) {
    const rt_image_params &ip = *c.ip;
    const rt_camera_uniform &cam = *c.cam;
    if (px >= ip.width || py >= ip.height) return;
    const uint32_t base_seed = py * ip.width + px + ip.frame * 2654435761u;
    Rng rng = rng_init(base_seed * 1664525u + 1013904223u);
    PixelAcc acc;
    acc.col = splat(0.0f);
    acc.lum_sum = 0.0f;
    acc.lum2_sum = 0.0f;
    for (int i = 0; i < 8; i++) acc.sample_lum[i] = 0.0f;
    V3 first_normal = v3(0.0f, 1.0f, 0.0f);
    float first_depth = 1e30f;
    const uint32_t spp = std::max(1u, ip.spp);
    const uint32_t debug_view = ip._pad[0];

    float env_intensity = 0.0f;
    V3 env_color = v3(0.55f, 0.6f, 0.75f);
    V3 env_horizon = v3(0.95f, 0.93f, 0.9f);
    for (uint32_t i = 0; i < c.s->n_lights; i++) {
        if (c.s->lights[i].kind == 3u) {
            env_intensity = c.s->lights[i].intensity;
            env_color = L3(c.s->lights[i].color);
            env_horizon = L3(c.s->lights[i].data);
            break;
        }
    }

    for (uint32_t s = 0u; s < spp; s++) {
        const float jx = rng_f32(&rng) - 0.5f;
        const float jy = rng_f32(&rng) - 0.5f;
        const V2 cam_inv = L2(cam.inv_half_res);
        const V2 uv = V2{((float)px + jx) * cam_inv.x - 1.0f, 1.0f - ((float)py + jy) * cam_inv.y};
        V3 ro = L3(cam.position);
        V3 rd = normalize(L3(cam.forward) + L3(cam.right) * uv.x + L3(cam.up) * uv.y);
        V3 thr = splat(1.0f);

        for (uint32_t b = 0u; b < c.max_depth; b++) {
            const Hit hit = trace_bvh(c.s, ro, rd, &c.st);
            if (hit.t >= INF) {
                if (b > 0u) add_radiance(s, acc, thr * sky_color(rd, env_color, env_horizon, env_intensity));
                break;
            }
            const rt_triangle &tr = c.s->tris[hit.prim_id];
            const rt_material &mat = c.s->mats[tr.material_id];
            const V3 hp = ro + rd * hit.t;
            const Bary bb = barycentric(tr, ro, rd, hit.t);
            const V2 uv_hit = bb.uv;

            if ((mat.flags & RT_MAT_FLAG_ALPHA_CUTOUT) != 0u && mat.tex_albedo != RT_TEX_NONE) {
                if (tex(c.s, RT_TEX_ROLE_ALBEDO, uv_hit, mat.tex_albedo).w < 0.5f) {  // gather.wgsl:318-323
                    ro = hp + rd * 0.002f;
                    continue;
                }
            }

            const V3 gn = get_triangle_normal(tr);
            const bool ff = dot(rd, gn) < 0.0f;
            const V3 sn_geom = smooth_normal(tr, bb, gn);
            const V3 sn = ff ? sn_geom : -sn_geom;
            V3 shn = sn;
            if (mat.tex_normal != RT_TEX_NONE) {  // gather.wgsl:335-340
                const V3 nm = tex(c.s, RT_TEX_ROLE_NORMAL, uv_hit, mat.tex_normal).rgb() * 2.0f - splat(1.0f);
                const TanFrame fr = tangent_frame(tr, sn);
                shn = normalize(fr.t * nm.x + fr.b * nm.y + sn * nm.z);
            }

            if (b == 0u && hit.t < first_depth) {
                first_normal = sn;
                first_depth = hit.t;
            }

            const MatParams params = eval_material(c.s, mat, uv_hit);
            const V3 albedo = params.albedo;
            const float rough = params.rough;
            const float metal = params.metal;
            const V3 emissive = emissive_at(c.s, mat, uv_hit);

            add_radiance(s, acc, thr * emissive);

            if (debug_view != 0u) {
                bool stop = true;
                switch (debug_view) {
                    case 1u: acc.col = v3(uv_hit.x, uv_hit.y, 0.0f); break;
                    case 2u: acc.col = sn * 0.5f + splat(0.5f); break;
                    case 3u: acc.col = shn * 0.5f + splat(0.5f); break;
                    case 4u: acc.col = gn * 0.5f + splat(0.5f); break;
                    case 5u: acc.col = emissive; break;
                    case 8u: case 9u: {  // screen-space samples of layer 0
                        const V2 puv = V2{(float)px / (float)ip.width, (float)py / (float)ip.height};
                        acc.col = tex(c.s, debug_view == 8u ? RT_TEX_ROLE_EMISSIVE : RT_TEX_ROLE_ALBEDO, puv, 0u).rgb();
                        break;
                    }
                    case 11u: acc.col = tex(c.s, RT_TEX_ROLE_EMISSIVE, uv_hit, 0u).rgb(); break;
                    case 12u: acc.col = L3(mat.emissive); break;
                    case 18u: acc.col = tex(c.s, RT_TEX_ROLE_ALBEDO, uv_hit, mat.tex_albedo).rgb(); break;
                    case 13u: acc.col = v3((float)mat.tex_emissive / 65535.0f, 0.0f, 0.0f); break;
                    case 17u:
                        acc.col = v3(uv_hit.x - std::floor(uv_hit.x), uv_hit.y - std::floor(uv_hit.y), 0.0f);
                        break;
                    case 19u: acc.col = tangent_frame(tr, sn).t * 0.5f + splat(0.5f); break;
                    case 20u: acc.col = tangent_frame(tr, sn).b * 0.5f + splat(0.5f); break;
                    case 21u: acc.col = hp; break;
                    default: stop = false; break;
                }
                if (stop) break;
            }

            if (mat.kind == 2u) {
                const V3 f0 = splat(wgsl_pow((mat.ior - 1.0f) / (mat.ior + 1.0f), 2.0f));
                const float ct = std::fabs(dot(rd, shn));
                const float fr = fresnel_schlick(ct, f0).x;
                if (rng_f32(&rng) < fr) {
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

            const V3 f0 = mix(splat(0.04f), albedo, splat(metal));
            const V3 wo = -rd;
            const float a = wmax(rough * rough, 1e-3f);
            const V3 fr = fresnel_schlick(std::fabs(dot(rd, shn)), f0);
            const V3 diffuse_weight = (1.0f - metal) * albedo;
            const float sp = vmax3(fr);
            const float dp = vmax3(diffuse_weight);
            const float sel = (sp + dp <= 0.0f) ? 0.5f : sp / (sp + dp);
            const float spec_prob = wclamp(sel, 0.001f, 0.999f);
            const bool do_spec = rng_f32(&rng) < spec_prob;
            const V3 dscale = diffuse_weight / (1.0f - spec_prob);

            if (!do_spec) {
                const LightSample ls = direct_light(c, hp, shn, &rng);
                if (env_intensity > 0.0f)
                    add_radiance(s, acc, thr * dscale * sky_color(shn, env_color, env_horizon, env_intensity));
                add_radiance(s, acc, thr * dscale * ls.e / PI);
                add_radiance(s, acc, thr * albedo * ip.ambient);
            } else {
                const LightSample ls = direct_light(c, hp, shn, &rng);
                add_radiance(s, acc, thr * spec_ggx(wo, shn, ls.wi, a, f0) * ls.e / spec_prob);
            }

            if (ip.use_photons == 1u && !do_spec) {
                uint32_t ph_hits = 0u;
                uint64_t reads = 0u;
                const V3 ph_sum = photon_estimate(c.pm, hp, shn, ip.photon_radius, &ph_hits, &reads);
                c.st.photon_reads += reads;
                if (ph_hits > 0u) {
                    add_radiance(s, acc,
                                 thr * dscale * ph_sum * ip.photon_scale /
                                     (PI * (float)ph_hits * ip.photon_radius * ip.photon_radius));
                }
                if (ip.photon_debug == 1u) {
                    const float density = wmin((float)ph_hits / 64.0f, 1.0f);
                    acc.col = v3(1.0f - density, density, 0.0f) * 2.0f;
                }
            }

            if (do_spec) {
                rd = sample_spec_dir(rd, shn, a, &rng);
                ro = hp + shn * 0.001f;
                const V3 h = normalize(wo + rd);
                const float ndl = wmax(dot(shn, rd), 0.0f);
                thr = thr * fresnel_schlick(wmax(dot(wo, h), 0.0f), f0) * smith_g1(ndl, a) / spec_prob;
            } else {
                const V3 local = sample_cosine_hemisphere(&rng);
                rd = to_world(local, shn);
                ro = hp + shn * 0.001f;
                thr = thr * dscale;
                if (ip.use_photons == 1u) break;
            }

            const float p = vmax3(thr);
            if (p <= 0.0f || rng_f32(&rng) > p) break;
            thr = thr / p;
        }
    }

    // gather.wgsl:548-579, firefly sigma clamp
    if (spp <= 8u) {
        float *sl = acc.sample_lum;
        for (uint32_t i = 1u; i < spp; i++) {
            const float key = sl[i];
            uint32_t j = i;
            while (j > 0u && sl[j - 1u] > key) {
                sl[j] = sl[j - 1u];
                j = j - 1u;
            }
            sl[j] = key;
        }
        const float median = (sl[(spp - 1u) / 2u] + sl[spp / 2u]) * 0.5f;
        const uint32_t half = std::max(spp / 2u, 1u);
        float spread = 0.0f;
        for (uint32_t i = 0u; i < half; i++) {
            const float d = sl[i] - median;
            spread += d * d;
        }
        const float sigma = wmax(wgsl_sqrt(spread / (float)half), 0.25f * median + 1e-5f);
        const float thresh = median + 3.0f * sigma;
        float clamped_lum = 0.0f, clamped_lum2 = 0.0f;
        for (uint32_t i = 0u; i < spp; i++) {
            const float lc = wmin(sl[i], thresh);
            clamped_lum += lc;
            clamped_lum2 += lc * lc;
        }
        acc.col = acc.col * (clamped_lum / wmax(acc.lum_sum, 1e-10f));
        acc.lum_sum = clamped_lum;
        acc.lum2_sum = clamped_lum2;
    }

    const V3 col = vmin(acc.col / (float)spp, splat(10.0f));
    const uint64_t idx = (uint64_t)py * ip.width + px;
    frame[4 * idx + 0] = col.x;
    frame[4 * idx + 1] = col.y;
    frame[4 * idx + 2] = col.z;
    frame[4 * idx + 3] = 1.0f;
    moments[4 * idx + 0] = wmin(acc.lum_sum / (float)spp, 10.0f);
    moments[4 * idx + 1] = wmin(acc.lum2_sum / (float)spp, 100.0f);
    moments[4 * idx + 2] = 0.0f;
    moments[4 * idx + 3] = 0.0f;
    gbuffer[4 * idx + 0] = first_normal.x;
    gbuffer[4 * idx + 1] = first_normal.y;
    gbuffer[4 * idx + 2] = first_normal.z;
    gbuffer[4 * idx + 3] = first_depth;
}

bool materials_supported(const or_scene *s) {
    for (uint32_t r = 0; r < RT_TEX_ROLES; r++)
        if (!s->tex[r].texels || s->tex[r].width == 0 || s->tex[r].height == 0 || s->tex[r].layers == 0) return false;
    return s->n_lights > 0;
}

// photon_trace.wgsl:91-97
inline V3 bounce_dir(V3 rd, V3 n, float rough, Rng *rng) {
    const V3 mirror = reflect(rd, n);
    const float blur = wgsl_smoothstep(0.0f, 1.0f, rough);
    if (blur <= 0.001f) return mirror;
    const V3 diff = sample_cosine_hemisphere(rng);
    return normalize(mix(mirror, to_world(diff, n), blur));
}

// photon_trace.wgsl:101-175: light pick and emission of photon `photon_index`
// (false: no light, or a light kind that emits nothing). Leaves the RNG where
// the trace continues.
static bool photon_emit(const or_scene *s, const rt_grid_params *gp, uint32_t photon_index, Rng &rng, V3 &origin,
                        V3 &dir, V3 &power, uint32_t *light_index) {
    const uint32_t num_lights = s->n_lights;
    if (num_lights == 0u) return false;
    rng = rng_init(photon_index * 1103515245u + 12345u);
    const uint32_t li = clamp_index(wgsl_u32(rng_f32(&rng) * (float)num_lights), num_lights, nullptr);
    if (light_index) *light_index = li;
    const rt_light &light = s->lights[li];
    const V3 ldata = L3(light.data);
    const V3 lcolor = L3(light.color);
    origin = ldata;
    if (light.kind == 0u) {
        float u1, u2;
        for (;;) {
            u1 = rng_f32(&rng) * 2.0f - 1.0f;
            u2 = rng_f32(&rng) * 2.0f - 1.0f;
            if (u1 * u1 + u2 * u2 < 1.0f) break;
        }
        const float z = 1.0f - 2.0f * (u1 * u1 + u2 * u2);
        const float r = wgsl_sqrt(1.0f - z * z);
        dir = normalize(v3(r * 2.0f * u1, r * 2.0f * u2, z));
        power = lcolor * light.intensity;
    } else if (light.kind == 1u) {
        const float disk_radius = 10.0f;
        const float disk_u = (rng_f32(&rng) - 0.5f) * disk_radius;
        const float disk_v = (rng_f32(&rng) - 0.5f) * disk_radius;
        const V3 up = (std::fabs(light.data[1]) < 0.999f) ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
        const V3 tangent = normalize(cross(up, ldata));
        const V3 bitangent = cross(ldata, tangent);
        origin = disk_u * tangent + disk_v * bitangent - ldata * 100.0f;
        dir = ldata;
        power = lcolor * light.intensity;
    } else if (light.kind == 2u) {
        const uint32_t ti = clamp_index(wgsl_u32(light.data[0]), s->n_tris, nullptr);
        const rt_triangle &tri = s->tris[ti];
        const float u1 = rng_f32(&rng);
        const float u2 = rng_f32(&rng);
        const float su = wgsl_sqrt(u1);
        const float u = 1.0f - su;
        const float v = u2 * su;
        const float w = 1.0f - u - v;
        origin = L3(tri.v0) * u + L3(tri.v1) * v + L3(tri.v2) * w;
        const V3 n = get_triangle_normal(tri);
        const V3 local_dir = sample_cosine_hemisphere(&rng);
        dir = to_world(local_dir, n);
        power = lcolor * light.intensity;
    } else if (light.kind == 3u) {
        const V3 local_dir = sample_cosine_hemisphere(&rng);
        const V3 d = to_world(local_dir, v3(0.0f, 1.0f, 0.0f));
        const V3 span = v3((float)gp->grid_x, (float)gp->grid_y, (float)gp->grid_z) * gp->cell_size;
        const V3 center = v3(gp->room_min_x, gp->room_min_y, gp->room_min_z) + span * 0.5f;
        const float R = wmax(wmax(span.x, span.y), span.z) * 1.2f + 0.05f;
        origin = center + d * R;
        dir = -d;
        const float t = wclamp(d.y * 0.5f + 0.5f, 0.0f, 1.0f);
        power = mix(lcolor, ldata, wgsl_pow(t, 0.5f)) * light.intensity;
    } else {
        return false;
    }
    return true;
}

// photon_trace.wgsl:101-260, one invocation.
void photon_one(const or_scene *s, const rt_grid_params *gp, uint32_t photon_index, uint32_t max_depth,
                rt_photon *photons) {
    Rng rng;
    V3 origin, dir, power;
    if (!photon_emit(s, gp, photon_index, rng, origin, dir, power, nullptr)) return;

    for (uint32_t bounce = 0u; bounce < max_depth; bounce++) {
        const Hit hit = trace_bvh(s, origin, dir, nullptr);
        if (hit.t >= INF) return;
        const rt_triangle &tri = s->tris[hit.prim_id];
        const rt_material &mat = s->mats[tri.material_id];
        const V3 hit_pos = origin + dir * hit.t;
        const Bary bb = barycentric(tri, origin, dir, hit.t);
        const V3 normal = get_triangle_normal(tri);
        const bool front_face = dot(dir, normal) < 0.0f;
        const V3 sn_geom = smooth_normal(tri, bb, normal);
        const V3 face_normal = front_face ? sn_geom : -sn_geom;
        V3 shn = face_normal;
        if (mat.tex_normal != RT_TEX_NONE) {  // photon_trace.wgsl:193-198
            const V3 nm = tex(s, RT_TEX_ROLE_NORMAL, bb.uv, mat.tex_normal).rgb() * 2.0f - splat(1.0f);
            const TanFrame fr = tangent_frame(tri, face_normal);
            shn = normalize(fr.t * nm.x + fr.b * nm.y + face_normal * nm.z);
        }

        if ((mat.flags & RT_MAT_FLAG_ALPHA_CUTOUT) != 0u && mat.tex_albedo != RT_TEX_NONE) {
            if (tex(s, RT_TEX_ROLE_ALBEDO, bb.uv, mat.tex_albedo).w < 0.5f) {  // photon_trace.wgsl:200-206
                origin = hit_pos + dir * 0.002f;
                continue;
            }
        }
        if (mat.tex_emissive != RT_TEX_NONE || length(L3(mat.emissive)) > 1e-4f || mat.kind == 3u) return;

        if (mat.kind == 2u) {
            const V3 f0 = splat(wgsl_pow((mat.ior - 1.0f) / (mat.ior + 1.0f), 2.0f));
            const float cos_theta = std::fabs(dot(dir, face_normal));
            const float fr = fresnel_schlick(cos_theta, f0).x;
            if (rng_f32(&rng) < fr) {
                dir = reflect(dir, face_normal);
                origin = hit_pos + face_normal * 0.001f;
            } else {
                const float eta = front_face ? 1.0f / mat.ior : mat.ior;
                dir = refract(dir, face_normal, eta);
                origin = hit_pos - face_normal * 0.001f;
            }
            continue;
        }

        rt_photon &ph = photons[(uint64_t)photon_index * max_depth + bounce];
        ph.position[0] = hit_pos.x; ph.position[1] = hit_pos.y; ph.position[2] = hit_pos.z;
        ph.power[0] = power.x; ph.power[1] = power.y; ph.power[2] = power.z;
        ph.incident[0] = -dir.x; ph.incident[1] = -dir.y; ph.incident[2] = -dir.z;
        ph.normal[0] = shn.x; ph.normal[1] = shn.y; ph.normal[2] = shn.z;

        // photon_trace.wgsl:84-89 surface_albedo: the texture alone when present
        const V3 albedo = mat.tex_albedo != RT_TEX_NONE ? tex(s, RT_TEX_ROLE_ALBEDO, bb.uv, mat.tex_albedo).rgb()
                                                        : v3(mat.albedo[0], mat.albedo[1], mat.albedo[2]);
        const V3 f0 = mix(splat(0.04f), albedo, splat(mat.metallic));
        const float ct = std::fabs(dot(dir, shn));
        const V3 fr = fresnel_schlick(ct, f0);
        const V3 diffuse_weight = (1.0f - mat.metallic) * albedo;
        const float spec_prob = wclamp(vmax3(fr), 0.0f, 0.98f);
        const bool do_spec = rng_f32(&rng) < spec_prob;
        if (do_spec) {
            dir = bounce_dir(dir, shn, mat.roughness, &rng);
            power = power * f0;
        } else {
            const V3 local_dir = sample_cosine_hemisphere(&rng);
            dir = to_world(local_dir, shn);
            power = power * diffuse_weight;
        }
        origin = hit_pos + shn * 0.001f;
        const float p = vmax3(power);
        if (p <= 0.0f || rng_f32(&rng) > p) return;
        power = power / p;
    }
}

// photon_count.wgsl:18-23 (u32 conversions saturate)
inline uint32_t photon_cell_u(const rt_grid_params *gp, const float *pos) {
    const uint32_t gx = wgsl_u32((pos[0] - gp->room_min_x) / gp->cell_size);
    const uint32_t gy = wgsl_u32((pos[1] - gp->room_min_y) / gp->cell_size);
    const uint32_t gz = wgsl_u32((pos[2] - gp->room_min_z) / gp->cell_size);
    return gx + gy * gp->grid_x + gz * gp->grid_x * gp->grid_y;
}

}  // namespace

extern "C" {

uint32_t or_rng_init(uint32_t seed) { return rng_init(seed).state; }

/* textureSampleLevel of one texture array (include/rt_texture.h). */
void or_tex_sample(const rt_tex_array *t, float u, float v, uint32_t layer, float *out) {
    rt_tex_sample(t, u, v, layer, out);
}
uint32_t or_rng_next(uint32_t *state) {
    Rng r{*state, 1u};
    const uint32_t w = rng_next(&r);
    *state = r.state;
    return w;
}
float or_rng_f32(uint32_t *state) {
    Rng r{*state, 1u};
    const float f = rng_f32(&r);
    *state = r.state;
    return f;
}
float or_intersect_triangle(const float *ro, const float *rd, const float *v0, const float *v1,
                            const float *v2) {
    return intersect_triangle(L3(ro), L3(rd), L3(v0), L3(v1), L3(v2));
}

void or_math_eval(int fn, const float *x, const float *y, float *out, uint32_t n) {
    for (uint32_t i = 0; i < n; i++) {
        switch (fn) {
            case 0: out[i] = wgsl_sin(x[i]); break;
            case 1: out[i] = wgsl_cos(x[i]); break;
            case 2: out[i] = wgsl_exp2(x[i]); break;
            case 3: out[i] = wgsl_log2(x[i]); break;
            case 4: out[i] = wgsl_exp(x[i]); break;
            case 5: out[i] = wgsl_pow(x[i], y[i]); break;
            case 6: out[i] = wgsl_sqrt(x[i]); break;
            case 7: out[i] = wgsl_inverse_sqrt(x[i]); break;
            case 8: out[i] = (float)wgsl_u32(x[i]); break;
            case 9: out[i] = (float)wgsl_i32(x[i]); break;
            case 10: out[i] = wgsl_smoothstep(0.0f, 1.0f, x[i]); break;
            default: out[i] = 0.0f; break;
        }
    }
}

int or_trace_closest(const or_scene *s, const rt_ray *rays, uint32_t n, rt_hit *out, or_stats *st) {
    for (uint32_t i = 0; i < n; i++) {
        const Hit h = trace_bvh(s, L3(rays[i].origin), L3(rays[i].dir), st);
        out[i].t = h.t;
        out[i].prim_id = h.prim_id;
    }
    return 0;
}

// Brute force over every BVH slot in slot order with the same strict `<`:
// the closest hit with ties going to the lowest slot, which is the first one
// met by a left-first depth-first traversal of the reference's BVH.
int or_trace_brute(const or_scene *s, const rt_ray *rays, uint32_t n, rt_hit *out) {
    for (uint32_t i = 0; i < n; i++) {
        const V3 ro = L3(rays[i].origin), rd = L3(rays[i].dir);
        float best = INF;
        uint32_t best_id = 0u;
        for (uint32_t k = 0; k < s->n_prims; k++) {
            const rt_triangle &tr = s->tris[s->prims[k]];
            const float t = intersect_triangle(ro, rd, L3(tr.v0), L3(tr.v1), L3(tr.v2));
            if (t < best) {
                best = t;
                best_id = s->prims[k];
            }
        }
        out[i].t = best;
        out[i].prim_id = best_id;
    }
    return 0;
}

int or_trace_brute_watertight(const or_scene *s, const rt_ray *rays, uint32_t n, rt_hit *out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        const V3 ro = L3(rays[i].origin), rd = L3(rays[i].dir);
        float best = INF;
        uint32_t best_id = 0u;
        for (uint32_t k = 0; k < s->n_tris; k++) {
            const rt_triangle &tr = s->tris[k];
            const float t = intersect_watertight(ro, rd, L3(tr.v0), L3(tr.v1), L3(tr.v2));
            if (t < best) {
                best = t;
                best_id = k;
            }
        }
        out[i].t = best;
        out[i].prim_id = best_id;
    }
    return 0;
}

int or_watertight_t(const or_scene *s, const rt_ray *rays, const uint32_t *prim_ids, uint32_t n, float *t_out) {
    for (uint32_t i = 0; i < n; i++) {
        if (prim_ids[i] >= s->n_tris) return -1;
        const rt_triangle &tr = s->tris[prim_ids[i]];
        t_out[i] = intersect_watertight(L3(rays[i].origin), L3(rays[i].dir), L3(tr.v0), L3(tr.v1), L3(tr.v2));
    }
    return 0;
}

int or_gather(const or_scene *s, const rt_camera_uniform *cam, const rt_image_params *ip,
              const or_photon_map *pm, uint32_t max_depth, uint32_t row_begin, uint32_t row_end,
              float *frame, float *moments, float *gbuffer, int n_threads, or_stats *st) {
    if (!materials_supported(s)) return -1;
    if (ip->use_photons == 1u && (pm == nullptr || pm->gp == nullptr || pm->grid_meta == nullptr)) return -2;
    row_end = std::min(row_end, ip->height);
    or_stats total{};
#pragma omp parallel num_threads(n_threads > 0 ? n_threads : 1)
    {
        Ctx c{s, ip, cam, pm, max_depth, {}};
#pragma omp for schedule(dynamic, 1)
        for (int64_t y = (int64_t)row_begin; y < (int64_t)row_end; y++) {
            for (uint32_t x = 0; x < ip->width; x++) gather_pixel(c, x, (uint32_t)y, frame, moments, gbuffer);
        }
#pragma omp critical
        {
            total.ray_queries += c.st.ray_queries;
            total.nodes_visited += c.st.nodes_visited;
            total.tris_tested += c.st.tris_tested;
            total.photon_reads += c.st.photon_reads;
            total.light_index_overflow += c.st.light_index_overflow;
        }
    }
    if (st) *st = total;
    return 0;
}

int or_photon_emit(const or_scene *s, const rt_grid_params *gp, uint32_t photon_index, float origin[3],
                   float dir[3], float power[3], uint32_t *light_index) {
    Rng rng;
    V3 o, d, pw;
    if (!photon_emit(s, gp, photon_index, rng, o, d, pw, light_index)) return 1;
    origin[0] = o.x, origin[1] = o.y, origin[2] = o.z;
    dir[0] = d.x, dir[1] = d.y, dir[2] = d.z;
    power[0] = pw.x, power[1] = pw.y, power[2] = pw.z;
    return 0;
}

int or_sample_spec_dirs(const float rd[3], const float n[3], float a, uint32_t seed, uint32_t count, float *out) {
    Rng rng = rng_init(seed);
    for (uint32_t i = 0; i < count; i++) {
        const V3 o = sample_spec_dir(v3(rd[0], rd[1], rd[2]), v3(n[0], n[1], n[2]), a, &rng);
        out[3 * i] = o.x, out[3 * i + 1] = o.y, out[3 * i + 2] = o.z;
    }
    return 0;
}

int or_photon_trace(const or_scene *s, const rt_grid_params *gp, uint32_t n_photons, uint32_t max_depth,
                    rt_photon *photons, int n_threads) {
    if (!materials_supported(s)) return -1;
#pragma omp parallel for schedule(dynamic, 256) num_threads(n_threads > 0 ? n_threads : 1)
    for (int64_t i = 0; i < (int64_t)n_photons; i++) photon_one(s, gp, (uint32_t)i, max_depth, photons);
    return 0;
}

// photon_count.wgsl:25-34, lib.rs:987-1007, photon_scatter.wgsl:27-37 (stable)
int or_photon_grid(const rt_photon *photons, uint32_t n_slots, const rt_grid_params *gp, uint32_t meta_cells,
                   uint32_t *grid_meta, rt_photon *sorted, uint32_t *n_sorted) {
    if (gp->num_cells > meta_cells) return -1;
    std::vector<uint32_t> counts(meta_cells, 0u);
    const uint32_t slots = std::min(n_slots, gp->photon_slots);
    for (uint32_t slot = 0; slot < slots; slot++) {
        const rt_photon &ph = photons[slot];
        if (ph.power[0] <= 0.0f) continue;
        const uint32_t cell = photon_cell_u(gp, ph.position);
        if (cell >= gp->num_cells) continue;
        counts[cell]++;
    }
    uint32_t acc = 0u;
    for (uint32_t i = 0; i < meta_cells; i++) {
        grid_meta[2 * i] = acc;
        grid_meta[2 * i + 1] = counts[i];
        acc += counts[i];
    }
    std::vector<uint32_t> cursor(meta_cells, 0u);
    for (uint32_t slot = 0; slot < slots; slot++) {
        const rt_photon &ph = photons[slot];
        if (ph.power[0] <= 0.0f) continue;
        const uint32_t cell = photon_cell_u(gp, ph.position);
        if (cell >= gp->num_cells) continue;
        const uint32_t pos = cursor[cell]++;
        sorted[grid_meta[2 * cell] + pos] = ph;
    }
    if (n_sorted) *n_sorted = acc;
    return 0;
}

// denoise.wgsl:51-124
int or_denoise_pass(const float *in, float *out, const float *gbuffer, const float *moments,
                    const rt_denoise_params *p) {
    static const float kernel_1d[3] = {0.375f, 0.25f, 0.0625f};
    const int32_t W = (int32_t)p->width, H = (int32_t)p->height;
    const int32_t step = (int32_t)p->step_size;
#pragma omp parallel for schedule(static)
    for (int32_t py = 0; py < H; py++) {
        for (int32_t px = 0; px < W; px++) {
            const uint64_t idx = (uint64_t)py * p->width + (uint32_t)px;
            const V3 center = L3(in + 4 * idx);
            const V3 n_center = L3(gbuffer + 4 * idx);
            const float d_center = gbuffer[4 * idx + 3];
            const float lum_center = luma_of(center);
            float var_acc = 0.0f, var_cnt = 0.0f;
            for (int32_t qy = -1; qy <= 1; qy++) {
                for (int32_t qx = -1; qx <= 1; qx++) {
                    const int32_t nx = px + qx, ny = py + qy;
                    if (nx < 0 || ny < 0 || nx >= W || ny >= H) continue;
                    const uint64_t q = (uint64_t)ny * p->width + (uint32_t)nx;
                    const float mx = moments[4 * q], my = moments[4 * q + 1];
                    var_acc += wmax(my - mx * mx, 0.0f);
                    var_cnt += 1.0f;
                }
            }
            const float var_frame = var_acc / wmax(var_cnt, 1.0f);
            const float ema_scale = p->alpha / wmax(2.0f - p->alpha, 1e-4f);
            const float var_c = var_frame * ema_scale;
            V3 sum = splat(0.0f);
            float total_weight = 0.0f;
            for (int32_t dy = -2; dy <= 2; dy++) {
                for (int32_t dx = -2; dx <= 2; dx++) {
                    const int32_t nx = px + dx * step, ny = py + dy * step;
                    if (nx < 0 || ny < 0 || nx >= W || ny >= H) continue;
                    const uint64_t nidx = (uint64_t)ny * p->width + (uint32_t)nx;
                    const V3 neighbor = L3(in + 4 * nidx);
                    const V3 n_neighbor = L3(gbuffer + 4 * nidx);
                    const float d_neighbor = gbuffer[4 * nidx + 3];
                    const float sw = kernel_1d[std::abs(dx)] * kernel_1d[std::abs(dy)];
                    const float nw = wgsl_pow(wmax(dot(n_center, n_neighbor), 0.0f), p->phi_normal);
                    const float dw =
                        wgsl_exp(-std::fabs(d_center - d_neighbor) / (p->phi_depth * wmax(d_center, 1e-4f)));
                    const float dl = luma_of(neighbor) - lum_center;
                    const float cw = wgsl_exp(-(dl * dl) / (var_c * p->lambda + 1e-6f));
                    const float w = sw * nw * dw * cw;
                    sum = sum + neighbor * w;
                    total_weight += w;
                }
            }
            V3 o = center;
            if (total_weight > 1e-6f) o = sum / total_weight;
            out[4 * idx + 0] = o.x;
            out[4 * idx + 1] = o.y;
            out[4 * idx + 2] = o.z;
            out[4 * idx + 3] = 1.0f;
        }
    }
    return 0;
}

// composite.wgsl:16-26: accumulation = mix(prev, fresh, alpha)
int or_composite(const float *frame, float *accum, uint32_t w, uint32_t h, float alpha) {
    const uint64_t n = (uint64_t)w * h * 4;
    for (uint64_t i = 0; i < n; i++) accum[i] = accum[i] * (1.0f - alpha) + frame[i] * alpha;
    return 0;
}

// lib.rs:865-885, in glam's f32 arithmetic: normalize_or_zero(v) = v * (1 / |v|)
// unless |v| is 0 or not finite; cross and dot in component order.
static V3 nz(V3 v) {
    const float len = std::sqrt((v.x * v.x + v.y * v.y) + v.z * v.z);
    const float r = 1.0f / len;
    if (!(std::isfinite(r) && r > 0.0f)) return v3(0.0f, 0.0f, 0.0f);
    return v * r;
}
static V3 crs(V3 a, V3 b) { return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
static float dt(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

int or_project_marker(const float cam_position[3], const float cam_look_at[3], const float cam_up[3],
                      float fov_degrees, float near_plane, uint32_t width, uint32_t height, const float pos[3],
                      float radius, float out[4]) {
    const V3 position = L3(cam_position), look = L3(cam_look_at), up0 = L3(cam_up);
    const V3 forward = nz(look - position);
    const V3 right = nz(crs(forward, up0));
    const V3 up = crs(right, forward);
    const V3 v = L3(pos) - position;
    const float z = dt(v, forward);
    out[0] = out[1] = out[2] = out[3] = 0.0f;
    if (z <= std::fmax(near_plane, 1e-4f)) return 0;
    const float fov_scale = std::tan(fov_degrees * (3.14159265358979323846f / 180.0f) * 0.5f);
    const float aspect = (float)width / (float)height;
    const float x_ndc = dt(v, right) / (z * fov_scale * aspect);
    const float y_ndc = dt(v, up) / (z * fov_scale);
    const float visible =
        (std::fabs(x_ndc) <= 1.5f && std::fabs(y_ndc) <= 1.5f) ? std::sqrt(dt(v, v)) : 0.0f;
    out[0] = (x_ndc * 0.5f + 0.5f) * (float)width;
    out[1] = (0.5f - y_ndc * 0.5f) * (float)height;
    const float radius_px = (radius / (z * fov_scale)) * ((float)height * 0.5f);
    out[2] = std::fmax(radius_px, 1.5f);
    out[3] = visible;
    return 0;
}

// tonemap.wgsl:19-50, then the rgba8unorm store (round to nearest, half up).
int or_tonemap(const float *accum, const float *gbuffer, const rt_tonemap_uniform *p, uint32_t w, uint32_t h,
               uint8_t *rgba) {
    const float inv_gamma = 1.0f / p->gamma;
    for (uint32_t y = 0; y < h; y++) {
        for (uint32_t x = 0; x < w; x++) {
            const uint64_t idx = (uint64_t)y * p->width + x;
            const V3 hdr = L3(accum + 4 * idx);
            const V3 exposed = hdr * p->exposure;
            const V3 mapped = exposed / (exposed + splat(1.0f));
            V3 color = vpow(mapped, inv_gamma);
            if (p->marker[3] > 0.0f) {
                const float dxp = ((float)x + 0.5f) - p->marker[0];
                const float dyp = ((float)y + 0.5f) - p->marker[1];
                const float d = wgsl_sqrt(dxp * dxp + dyp * dyp);
                const float r = wmax(p->marker[2], 1.5f);
                const float hit_dist = gbuffer[4 * idx + 3];
                if (hit_dist > p->marker[3] - 0.2f) {
                    const float ball = 1.0f - wgsl_smoothstep(r * 0.75f, r, d);
                    const float halo = 1.0f - wgsl_smoothstep(r, r * 3.5f, d);
                    const V3 core = v3(1.0f, 0.94f, 0.80f) * wgsl_pow(ball, 1.4f) * 0.85f;
                    const V3 ring = v3(1.0f, 0.46f, 0.14f) * halo * halo * 0.8f;
                    color = color + (core + ring) * p->glow * (splat(1.0f) - 0.55f * color);
                }
            }
            const float cc[4] = {color.x, color.y, color.z, 1.0f};
            for (int k = 0; k < 4; k++) {
                const float v = wclamp(cc[k], 0.0f, 1.0f);
                rgba[4 * ((uint64_t)y * w + x) + k] = (uint8_t)(uint32_t)std::floor(v * 255.0f + 0.5f);
            }
        }
    }
    return 0;
}

// examples/common/mod.rs:37-47: (v/(1+v)).clamp(0,1).powf(1/2.2)*255 as u8, alpha 255
int or_tonemap_cpu(const float *accum, uint64_t n_px, uint8_t *rgba) {
    const float e = 1.0f / 2.2f;
    for (uint64_t i = 0; i < n_px; i++) {
        for (int k = 0; k < 3; k++) {
            const float v = accum[4 * i + k];
            float m = v / (1.0f + v);
            m = (m != m) ? m : std::min(std::max(m, 0.0f), 1.0f);
            const float o = std::pow(m, e) * 255.0f;
            rgba[4 * i + k] = (o != o) ? 0 : (o >= 255.0f ? 255 : (o <= 0.0f ? 0 : (uint8_t)o));
        }
        rgba[4 * i + 3] = 255;
    }
    return 0;
}


// fullscreen.wgsl:13-25: vs_main's fullscreen triangle gives each surface pixel
// uv = ((x + 0.5) / sw, (y + 0.5) / sh); fs_main samples the output texture
// there with the quad sampler (lib.rs:555-562: ClampToEdge, linear). The
// sampler's arithmetic is the shared rt_present_axis / rt_present_blend.
int or_present(const uint8_t *rgba, uint32_t tw, uint32_t th, uint8_t *bgra, uint32_t sw, uint32_t sh) {
    if (!rgba || !bgra || tw == 0 || th == 0) return -1;
    const float scale_x = rt_present_scale(tw, sw), scale_y = rt_present_scale(th, sh);
    for (uint32_t y = 0; y < sh; y++) {
        uint32_t y0, y1, ay;
        rt_present_axis(y, scale_y, th, &y0, &y1, &ay);
        for (uint32_t x = 0; x < sw; x++) {
            uint32_t x0, x1, ax;
            rt_present_axis(x, scale_x, tw, &x0, &x1, &ax);
            uint32_t t[4];
            const uint64_t at[4] = {(uint64_t)y0 * tw + x0, (uint64_t)y0 * tw + x1, (uint64_t)y1 * tw + x0,
                                    (uint64_t)y1 * tw + x1};
            for (int k = 0; k < 4; k++) {
                const uint8_t *p = rgba + 4 * at[k];
                t[k] = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
            }
            const uint32_t o = rt_present_blend(t[0], t[1], t[2], t[3], ax, ay);
            uint8_t *q = bgra + 4 * ((uint64_t)y * sw + x);
            for (int k = 0; k < 4; k++) q[k] = (uint8_t)(o >> (8 * k));
        }
    }
    return 0;
}

int or_photon_estimate(const or_photon_map *pm, const float hp[3], const float n[3], float radius, float sum[3],
                       uint32_t *hits, uint64_t *reads) {
    uint64_t r = 0;
    const V3 v = photon_estimate(pm, L3(hp), L3(n), radius, hits, &r);
    sum[0] = v.x;
    sum[1] = v.y;
    sum[2] = v.z;
    if (reads) *reads = r;
    return 0;
}

}  // extern "C"
