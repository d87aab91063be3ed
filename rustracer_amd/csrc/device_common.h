// device_common.h — device-side vector math, PCG, Möller–Trumbore and BVH
// traversal for the gfx950 kernels.
//
// Arithmetic follows WGSL's definitions in float32 with no FMA contraction
// (every kernel is compiled with -ffp-contract=off): dot = (x*x' + y*y') + z*z',
// normalize(v) = v * (1/sqrt(dot(v,v))), vector/scalar divides componentwise,
// and the implementation-defined builtins come from rt_wgsl_math.h. That keeps
// every per-pixel result identical to a CPU evaluation of the same WGSL.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rt_texture.h"
#include "../../include/rt_types.h"
#include "../../include/rt_wgsl_math.h"
#include "../../include/rustracer_hip.h"

namespace rtd {

constexpr float kPi = 3.141592653589793f;  // common.wgsl:8
constexpr float kInf = 1e38f;              // common.wgsl:9
constexpr float kEps = 1e-6f;              // common.wgsl:10

struct F3 {
    float x, y, z;
};
struct F2 {
    float x, y;
};

__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 fs(float s) { return F3{s, s, s}; }
__device__ __forceinline__ F3 ld3(const float *p) { return F3{p[0], p[1], p[2]}; }
__device__ __forceinline__ F3 xyz(float4 v) { return F3{v.x, v.y, v.z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 operator*(F3 a, F3 b) { return F3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ F3 operator/(F3 a, F3 b) { return F3{a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 operator*(float s, F3 a) { return F3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ F3 operator/(F3 a, float s) { return F3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ F3 operator-(F3 a) { return F3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
    return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(F3 v) { return wgsl_sqrt(dot(v, v)); }
__device__ __forceinline__ F3 normalize(F3 v) { return v * (1.0f / wgsl_sqrt(dot(v, v))); }
__device__ __forceinline__ F3 vmin(F3 a, F3 b) { return F3{wmin(a.x, b.x), wmin(a.y, b.y), wmin(a.z, b.z)}; }
__device__ __forceinline__ F3 mix(F3 a, F3 b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ F3 mix(F3 a, F3 b, F3 t) { return a * (fs(1.0f) - t) + b * t; }
__device__ __forceinline__ float vmax3(F3 v) { return wmax(wmax(v.x, v.y), v.z); }

// common.wgsl:12-25
__device__ __forceinline__ F3 reflect(F3 v, F3 n) { return v - (2.0f * dot(v, n)) * n; }
__device__ __forceinline__ F3 refract(F3 v, F3 n, float eta) {
    const float cos_theta = wmin(dot(-v, n), 1.0f);
    const F3 perp = eta * (v + cos_theta * n);
    const F3 par = (-wgsl_sqrt(__builtin_fabsf(1.0f - dot(perp, perp)))) * n;
    return perp + par;
}
__device__ __forceinline__ F3 fresnel_schlick(float cos_theta, F3 f0) {
    return f0 + (fs(1.0f) - f0) * wgsl_pow(wclamp(1.0f - cos_theta, 0.0f, 1.0f), 5.0f);
}

// PCG of common.wgsl:29-56 (the increment is always 1)
struct Rng {
    uint32_t state;
};
__device__ __forceinline__ Rng rng_init(uint32_t seed) {
    uint32_t s = seed * 747796405u + 2891336453u;
    s = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return Rng{s};
}
__device__ __forceinline__ uint32_t rng_next(Rng &r) {
    const uint32_t old = r.state;
    r.state = old * 747796405u + 1u;
    const uint32_t word = ((old >> ((old >> 28u) + 4u)) ^ old) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __forceinline__ float rng_f32(Rng &r) { return (float)rng_next(r) * 2.3283064365386963e-10f; }

// common.wgsl:60-74
__device__ __forceinline__ F3 sample_cosine_hemisphere(Rng &r) {
    const float u1 = rng_f32(r);
    const float u2 = rng_f32(r);
    const float phi = (2.0f * kPi) * u1;
    const float ct = wgsl_sqrt(u2);
    const float st = wgsl_sqrt(1.0f - u2);
    return F3{wgsl_cos(phi) * st, ct, wgsl_sin(phi) * st};
}
__device__ __forceinline__ F3 to_world(F3 l, F3 n) {
    const F3 up = (__builtin_fabsf(n.y) < 0.999f) ? f3(0.0f, 1.0f, 0.0f) : f3(1.0f, 0.0f, 0.0f);
    const F3 t = normalize(cross(up, n));
    const F3 b = cross(n, t);
    return t * l.x + n * l.y + b * l.z;
}

__device__ __forceinline__ F3 get_triangle_normal(const rt_triangle &tri) {
    const F3 v0 = ld3(tri.v0);
    return normalize(cross(ld3(tri.v1) - v0, ld3(tri.v2) - v0));
}

struct Bary {
    F2 uv;
    float u, v;
};

// textureSampleLevel(role, sampler, uv, i32(layer), 0.0), include/rt_texture.h.
struct F4s {
    float x, y, z, w;
};
__device__ __forceinline__ F4s tex_sample(const rt_tex_array &t, F2 uv, uint32_t layer) {
    float o[4];
    rt_tex_sample(&t, uv.x, uv.y, layer, o);
    return F4s{o[0], o[1], o[2], o[3]};
}
__device__ __forceinline__ F3 rgb(F4s c) { return F3{c.x, c.y, c.z}; }
// common.wgsl:178-192
__device__ __forceinline__ Bary barycentric(const rt_triangle &tri, F3 ro, F3 rd, float t) {
    const F3 v0 = ld3(tri.v0);
    const F3 e1 = ld3(tri.v1) - v0;
    const F3 e2 = ld3(tri.v2) - v0;
    const F3 p = (ro + rd * t) - v0;
    const float d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2);
    const float d20 = dot(p, e1), d21 = dot(p, e2);
    const float denom = d00 * d11 - d01 * d01;
    float u = 0.0f, v = 0.0f;
    if (denom != 0.0f) {
        u = (d11 * d20 - d01 * d21) / denom;
        v = (d00 * d21 - d01 * d20) / denom;
    }
    const float w = 1.0f - u - v;
    Bary b;
    b.uv = F2{(tri.uv0[0] * w + tri.uv1[0] * u) + tri.uv2[0] * v, (tri.uv0[1] * w + tri.uv1[1] * u) + tri.uv2[1] * v};
    b.u = u;
    b.v = v;
    return b;
}
// common.wgsl:196-201
__device__ __forceinline__ F3 smooth_normal(const rt_triangle &tri, const Bary &b, F3 geometric) {
    const float w = 1.0f - b.u - b.v;
    const F3 n = ld3(tri.n0) * w + ld3(tri.n1) * b.u + ld3(tri.n2) * b.v;
    if (dot(n, n) < 1e-8f) return geometric;
    return normalize(n);
}

// common.wgsl:211-230 (used by the tangent/bitangent debug views)
struct TanFrame {
    F3 t, b;
};
__device__ __forceinline__ TanFrame tangent_frame(const rt_triangle &tri, F3 n) {
    const F3 v0 = ld3(tri.v0);
    const F3 e1 = ld3(tri.v1) - v0;
    const F3 e2 = ld3(tri.v2) - v0;
    const F2 d1 = F2{tri.uv1[0] - tri.uv0[0], tri.uv1[1] - tri.uv0[1]};
    const F2 d2 = F2{tri.uv2[0] - tri.uv0[0], tri.uv2[1] - tri.uv0[1]};
    const float det = d1.x * d2.y - d2.x * d1.y;
    const float f = (__builtin_fabsf(det) < 1e-12f) ? 0.0f : 1.0f / det;
    const F3 t = (e1 * d2.y - e2 * d1.y) * f;
    const F3 b = (e2 * d1.x - e1 * d2.x) * f;
    F3 tt = t - n * dot(n, t);
    if (dot(tt, tt) < 1e-10f) {
        tt = cross(n, f3(1.0f, 0.0f, 0.0f));
        if (dot(tt, tt) < 1e-6f) tt = cross(n, f3(0.0f, 1.0f, 0.0f));
    }
    tt = normalize(tt);
    const F3 bb = cross(n, tt) * ((dot(cross(n, tt), b) > 0.0f) ? 1.0f : -1.0f);
    return TanFrame{tt, bb};
}

// ─────────────────────────── traversal ───────────────────────────

// Slot-ordered traversal triangles: three float4 per BVH slot:
// (v0, prim index bits), (v1 - v0, 0), (v2 - v0, 0). The edges are the same
// float subtractions Möller–Trumbore performs, so results are unchanged.
struct TraceScene {
    const rt_bvh_node *nodes;  // reference layout, root at 0
    const uint32_t *prims;     // slot -> primitive index
    const float4 *slot_tris;   // 3 float4 per slot
    const float4 *wnodes;      // child-box layout: 4 float4 per internal node
    const float4 *w4nodes;     // 4-wide layout: 8 float4 per record
    uint32_t n_nodes;
    uint32_t pad;
    float4 root_min;           // root box min (xyz) and root ref bits (w) for the child-box layout
    float4 root_max;           // root box max (xyz)
};

struct Hit {
    float t;
    uint32_t prim;  // primitive index (closest hit)
};

// Counters of the counting build. Phases: traversal steps of closest-hit
// queries, one bounce body, one batch of photon candidates, traversal steps of
// shadow queries, one iteration of the persistent camera pass's loop (its live
// lanes). Per phase: lane steps and wave steps (their ratio is the mean number
// of active lanes) and the shader-clock cycles wavefronts spent in the
// phase's code regions (each execution of a region by a wavefront counted
// once; for the loop phase, whole iterations).
enum CountPhase { kPhaseTraversal = 0, kPhaseBounce, kPhasePhotons, kPhaseShadow, kPhaseLoop, kCountPhases };
static_assert(kCountPhases == RT_COUNT_PHASES, "counter layout of rt_trace_counters");
// Counter words, in the order of rt_trace_counters.
enum CountWord {
    kCntQueries = 0,
    kCntNodes,
    kCntTris,
    kCntPhotons,
    kCntLaneSteps,
    kCntWaveSteps = kCntLaneSteps + kCountPhases,
    kCntCycles = kCntWaveSteps + kCountPhases,
    kCntWords = kCntCycles + kCountPhases
};
typedef __attribute__((address_space(3))) unsigned long long lds_u64;

__device__ __forceinline__ bool first_active_lane() {
    const unsigned long long active = __ballot(1);
    return __lane_id() == (uint32_t)(__ffsll((long long)active) - 1);
}
// The counters live in LDS, one set per wavefront (the counting kernels run
// one-wavefront workgroups): each update adds the wavefront's total, with an
// LDS atomic issued by its first active lane (an atomic, so the compiler keeps
// no per-lane copy of a counter in a register across a loop), so counting
// keeps nothing in registers.
__device__ __forceinline__ void lds_add(lds_u64 *p, unsigned long long v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
struct TraceCount {
    lds_u64 *w;           // kCntWords words (null outside the counting build)
    uint32_t trav_phase;  // the phase the traversal loop counts into
    // every active lane adds k to `word`
    __device__ __forceinline__ void add(int word, uint32_t k) {
        const unsigned long long active = __ballot(1);
        if (__lane_id() == (uint32_t)(__ffsll((long long)active) - 1))
            lds_add(&w[word], (unsigned long long)k * (unsigned long long)__popcll(active));
    }
};
__device__ __forceinline__ void count_step(TraceCount &c, int phase) {
    const unsigned long long active = __ballot(1);
    if (__lane_id() == (uint32_t)(__ffsll((long long)active) - 1)) {
        lds_add(&c.w[kCntLaneSteps + phase], (unsigned long long)__popcll(active));
        lds_add(&c.w[kCntWaveSteps + phase], 1ull);
    }
}
// A traversal loop step, counted into the phase of the query being traced.
__device__ __forceinline__ void count_traversal_step(TraceCount &c) {
    if (c.trav_phase == kPhaseShadow) count_step(c, kPhaseShadow);
    else count_step(c, kPhaseTraversal);
}
// The shader clock's low 32 bits (differences are taken modulo 2^32, exact
// for regions shorter than ~1.7 s).
__device__ __forceinline__ uint32_t phase_clock() { return (uint32_t)clock64(); }
// Time of a region of phase `phase` that started at t0 = phase_clock().
__device__ __forceinline__ void count_time(TraceCount &c, int phase, uint32_t t0) {
    const uint32_t dt = phase_clock() - t0;
    if (first_active_lane()) lds_add(&c.w[kCntCycles + phase], dt);
}
// Counting kernels: the LDS words zeroed at the start, added to the launch's
// counters at the end (by the first active lane; lanes may have left).
__device__ __forceinline__ void count_begin(unsigned long long *words) {
    if (first_active_lane())
        for (int k = 0; k < kCntWords; k++) ((lds_u64 *)words)[k] = 0ull;
}
__device__ __forceinline__ void count_flush(unsigned long long *counters, unsigned long long *words) {
    if (first_active_lane())
        for (int k = 0; k < kCntWords; k++) atomicAdd(&counters[k], ((lds_u64 *)words)[k]);
}

// Möller–Trumbore on a slot record, common.wgsl:78-100 (edges precomputed).
// Every quantity is computed and the shader's early returns become one
// select: the same value for every input (NaNs included), without branches
// (in a wavefront some lane nearly always gets past each test anyway).
__device__ __forceinline__ float mt_slot(F3 ro, F3 rd, F3 v0, F3 edge1, F3 edge2) {
    const F3 h = cross(rd, edge2);
    const float a = dot(edge1, h);
    const float f = 1.0f / a;
    const F3 s = ro - v0;
    const float u = f * dot(s, h);
    const F3 q = cross(s, edge1);
    const float v = f * dot(rd, q);
    const float t = f * dot(edge2, q);
    const bool hit = !(__builtin_fabsf(a) < 1e-7f) & !(u < 0.0f) & !(u > 1.0f) & !(v < 0.0f) & !(u + v > 1.0f) &
                     (t > kEps);
    return hit ? t : kInf;
}

// Watertight ray/triangle test (Woop, Benthin and Wald, "Watertight
// Ray/Triangle Intersection", JCGT 2(1), 2013), the opt-in alternative to
// Moller-Trumbore. The ray is set up once: the axis kz of its largest
// direction component, the other two axes (swapped when dir[kz] < 0, which
// keeps the winding) and the shear constants. Per triangle the vertices are
// taken relative to the origin, sheared so the ray runs along +z, and the
// three 2-D edge functions are evaluated; an edge shared by two triangles gives
// the same two products in both (the vertices are stored as they are, not as
// edges), so its functions are exact negatives and no ray passes between the
// triangles. An edge function that rounds to exactly zero is recomputed in
// double precision. Two-sided, with the reference's t > 1e-6.
struct WtRay {
    int kx, ky, kz;
    float sx, sy, sz;
};
__device__ __forceinline__ float comp(F3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
__device__ __forceinline__ WtRay wt_ray(F3 rd) {
    const float ax = __builtin_fabsf(rd.x), ay = __builtin_fabsf(rd.y), az = __builtin_fabsf(rd.z);
    WtRay w;
    w.kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    w.kx = w.kz == 2 ? 0 : w.kz + 1;
    w.ky = w.kx == 2 ? 0 : w.kx + 1;
    const float dz = comp(rd, w.kz);
    if (dz < 0.0f) {
        const int t = w.kx;
        w.kx = w.ky;
        w.ky = t;
    }
    w.sx = comp(rd, w.kx) / dz;
    w.sy = comp(rd, w.ky) / dz;
    w.sz = 1.0f / dz;
    return w;
}
__device__ __forceinline__ float wt_slot(const WtRay &w, F3 ro, F3 v0, F3 v1, F3 v2) {
    const F3 A = v0 - ro, B = v1 - ro, C = v2 - ro;
    const float az = comp(A, w.kz), bz = comp(B, w.kz), cz = comp(C, w.kz);
    const float ax = comp(A, w.kx) - w.sx * az, ay = comp(A, w.ky) - w.sy * az;
    const float bx = comp(B, w.kx) - w.sx * bz, by = comp(B, w.ky) - w.sy * bz;
    const float cx = comp(C, w.kx) - w.sx * cz, cy = comp(C, w.ky) - w.sy * cz;
    float u = cx * by - cy * bx;
    float v = ax * cy - ay * cx;
    float e = bx * ay - by * ax;
    if (u == 0.0f || v == 0.0f || e == 0.0f) {
        u = (float)((double)cx * (double)by - (double)cy * (double)bx);
        v = (float)((double)ax * (double)cy - (double)ay * (double)cx);
        e = (float)((double)bx * (double)ay - (double)by * (double)ax);
    }
    if ((u < 0.0f || v < 0.0f || e < 0.0f) && (u > 0.0f || v > 0.0f || e > 0.0f)) return kInf;
    const float det = u + v + e;
    if (det == 0.0f) return kInf;
    const float T = (u * (w.sz * az) + v * (w.sz * bz)) + e * (w.sz * cz);
    const float t = T / det;
    return t > kEps ? t : kInf;
}

// Per-lane traversal stack of words: the first kLdsStack words live in LDS,
// deeper words in a private array owned by the kernel (scratch memory). The
// LDS part is laid out in pairs: words 2j and 2j+1 of a lane are adjacent,
// and pair row j holds the pairs of the 64 lanes of the wavefront side by
// side, so a two-word entry (record, entry distance) is one 64-bit access and
// a row access by the wavefront covers 512 contiguous bytes. The spill array
// is only touched past kLdsStack words.
constexpr int kLdsStack = 32;
constexpr int kStackWords = 96;
constexpr int kSpillWords = kStackWords - kLdsStack;
constexpr int kSpillAlloc = kSpillWords > 0 ? kSpillWords : 1;
constexpr uint32_t kStackLanes = 64;  // lanes sharing the LDS rows (one wavefront per workgroup)

// Address-space-qualified pointers, so stack accesses compile to ds_* (LDS)
// and scratch_* (private) instructions instead of flat ones.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) u32x2 lds_u2;
typedef __attribute__((address_space(5))) uint32_t priv_u32;

struct LaneStack {
    lds_u32 *lds;     // this lane's pair 0 in the workgroup's kLdsStack * 64 LDS words
    priv_u32 *spill;  // kSpillAlloc private words
    __device__ __forceinline__ void init(uint32_t *rows, uint32_t *spill_words) {
        lds = (lds_u32 *)&rows[2u * (threadIdx.x % kStackLanes)];
        spill = (priv_u32 *)spill_words;
    }
    __device__ __forceinline__ static uint32_t at(uint32_t w) { return (w >> 1) * (2u * kStackLanes) + (w & 1u); }
    __device__ __forceinline__ void put(uint32_t sp, uint32_t v) {
        if (kSpillWords == 0 || sp < (uint32_t)kLdsStack) lds[at(sp)] = v;
        else spill[sp - kLdsStack] = v;
    }
    __device__ __forceinline__ uint32_t get(uint32_t sp) const {
        return (kSpillWords == 0 || sp < (uint32_t)kLdsStack) ? lds[at(sp)] : spill[sp - kLdsStack];
    }
    // Two-word entries at an even word index (pair row sp / 2).
    __device__ __forceinline__ static uint32_t at2(uint32_t sp) { return sp * kStackLanes; }
    __device__ __forceinline__ void put2_lds(uint32_t sp, uint32_t a, uint32_t b) {
        u32x2 v;
        v.x = a;
        v.y = b;
        *(lds_u2 *)&lds[at2(sp)] = v;
    }
    __device__ __forceinline__ uint2 get2_lds(uint32_t sp) const {
        const u32x2 v = *(const lds_u2 *)&lds[at2(sp)];
        return make_uint2(v.x, v.y);
    }
    __device__ __forceinline__ void put2(uint32_t sp, uint32_t a, uint32_t b) {
        if (kSpillWords == 0 || sp < (uint32_t)kLdsStack) {
            put2_lds(sp, a, b);
        } else {
            spill[sp - kLdsStack] = a;
            spill[sp + 1u - kLdsStack] = b;
        }
    }
    __device__ __forceinline__ uint2 get2(uint32_t sp) const {
        if (kSpillWords == 0 || sp < (uint32_t)kLdsStack) return get2_lds(sp);
        return make_uint2(spill[sp - kLdsStack], spill[sp + 1u - kLdsStack]);
    }
};

// Slab test of gather.wgsl:72-75; min/max on the GPU use v_min/v_max whose
// only difference from wmin/wmax is the sign of a zero, which no comparison sees.
__device__ __forceinline__ void slab(F3 bmin, F3 bmax, F3 ro, F3 inv, float &te, float &ta) {
    const float t0x = (bmin.x - ro.x) * inv.x, t1x = (bmax.x - ro.x) * inv.x;
    const float t0y = (bmin.y - ro.y) * inv.y, t1y = (bmax.y - ro.y) * inv.y;
    const float t0z = (bmin.z - ro.z) * inv.z, t1z = (bmax.z - ro.z) * inv.z;
    te = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    ta = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
}

// Conservative slab test of the watertight mode (Ize, "Robust BVH Ray
// Traversal", JCGT 2(2), 2013): an axis whose product is 0 * inf (a zero
// direction component with the origin in a face plane) is inside the slab for
// every t, and the interval is widened at both ends by kWtBoxMargin of its
// magnitude, far beyond the 2*gamma_3 that the rounding of (b - o) * (1 / d)
// can take away, plus `slack`. The slack covers the watertight test's own t:
// it interpolates the vertices' distances along the ray, so its error is
// absolute, a few ulps of the distance to the farthest vertex, not of t (a hit
// 1e-4 from the origin of a triangle 1 away is good to ~1e-7, not 1e-11).
// With slack = kWtBoxMargin * 2 * (L-inf distance from the origin to the far
// corner of the scene box), a box holding a hit nearer than the best so far is
// never culled, so traversal finds the hit a test of every triangle finds.
constexpr float kWtBoxMargin = 1.0f / 65536.0f;
__device__ __forceinline__ float wt_slack(F3 ro, F3 root_min, F3 root_max) {
    const float dx = fmaxf(__builtin_fabsf(root_min.x - ro.x), __builtin_fabsf(root_max.x - ro.x));
    const float dy = fmaxf(__builtin_fabsf(root_min.y - ro.y), __builtin_fabsf(root_max.y - ro.y));
    const float dz = fmaxf(__builtin_fabsf(root_min.z - ro.z), __builtin_fabsf(root_max.z - ro.z));
    return kWtBoxMargin * 2.0f * fmaxf(fmaxf(dx, dy), dz);
}
__device__ __forceinline__ void slab_axis_wt(float b0, float b1, float o, float inv, float &n, float &f) {
    const float t0 = (b0 - o) * inv, t1 = (b1 - o) * inv;
    n = fminf(t0, t1);
    f = fmaxf(t0, t1);
    if (t0 != t0 || t1 != t1) {
        n = -kInf;
        f = kInf;
    }
}
__device__ __forceinline__ void slab_wt(F3 bmin, F3 bmax, F3 ro, F3 inv, float slack, float &te, float &ta) {
    float nx, fx, ny, fy, nz, fz;
    slab_axis_wt(bmin.x, bmax.x, ro.x, inv.x, nx, fx);
    slab_axis_wt(bmin.y, bmax.y, ro.y, inv.y, ny, fy);
    slab_axis_wt(bmin.z, bmax.z, ro.z, inv.z, nz, fz);
    te = fmaxf(fmaxf(nx, ny), nz);
    ta = fminf(fminf(fx, fy), fz);
    te -= __builtin_fabsf(te) * kWtBoxMargin + slack;
    ta += __builtin_fabsf(ta) * kWtBoxMargin + slack;
}

// Reference-order traversal, gather.wgsl:66-88: left child first, strict
// `te < hit.t` and `t < hit.t`, right child dropped when 63 entries are used.
// EARLY_EXIT (shadow queries): the same traversal, stopped as soon as the best
// hit is nearer than t_max. Up to that point it is the reference's closest-hit
// traversal step for step, so "closest hit < t_max" is answered exactly.
template <bool EARLY_EXIT, bool COUNT>
__device__ Hit trace_reference(const TraceScene &S, F3 ro, F3 rd, float t_max, LaneStack &stk, TraceCount &cnt) {
    Hit hit{kInf, 0u};
    const F3 inv = fs(1.0f) / rd;
    uint32_t idx = 0u, sp = 0u;
    if (S.n_nodes == 0u) return hit;
    for (;;) {
        if (COUNT) count_traversal_step(cnt);
        const rt_bvh_node &n = S.nodes[idx];
        const float4 a = *reinterpret_cast<const float4 *>(&n.bbox_min[0]);
        const float4 b = *reinterpret_cast<const float4 *>(&n.bbox_max[0]);
        const uint32_t prim_count = n.prim_count;
        if (COUNT) cnt.add(kCntNodes, 1);
        float te, ta;
        slab(xyz(a), xyz(b), ro, inv, te, ta);
        bool pop = true;
        if (te <= ta && ta > 0.0f && te < hit.t) {
            const uint32_t left_first = __float_as_uint(a.w);
            if (prim_count > 0u) {
                for (uint32_t i = 0u; i < prim_count; i++) {
                    const uint32_t slot = left_first + i;
                    const float4 q0 = S.slot_tris[3 * slot], q1 = S.slot_tris[3 * slot + 1], q2 = S.slot_tris[3 * slot + 2];
                    if (COUNT) cnt.add(kCntTris, 1);
                    const float t = mt_slot(ro, rd, xyz(q0), xyz(q1), xyz(q2));
                    if (t < hit.t) {
                        hit.t = t;
                        hit.prim = __float_as_uint(q0.w);
                    }
                }
                if (EARLY_EXIT && hit.t < t_max) return hit;
            } else {
                if (sp < 63u) {
                    stk.put(sp, __float_as_uint(b.w));
                    sp += 1u;
                }
                idx = left_first;
                pop = false;
            }
        }
        if (pop) {
            if (sp == 0u) break;
            sp -= 1u;
            idx = stk.get(sp);
        }
    }
    return hit;
}

// Child-box traversal over the wide layout: each internal record holds both
// children's boxes, so one 64-byte load tests two boxes.
//
// Default (NEAREST = false): the reference's order. The left child is entered
// first; the right child's entry distance is kept on the stack and re-checked
// against the best hit when popped, which is exactly when the reference would
// test its box (gather.wgsl:75). Testing a child's box at its parent instead of
// at the child changes no comparison (same operands), and a box that fails
// `te < hit.t` now fails it later too. Together with the strict `t < hit.t` this
// returns the reference's hit bit for bit, including its order-dependent
// choice when a triangle's rounded t lands just before its own box's entry.
// EARLY_EXIT: shadow query, stopped once the best hit is nearer than t_max.
//
// NEAREST = true: the nearer child first and lexicographic (t, slot) ties. It
// equals the reference except where a hit rounds to before its box's entry
// (measured, not guaranteed; opt-in via RT_FLAG_NEAREST_FIRST).
//
// Record i (internal nodes only, root = 0):
//   {lmin.xyz, lref}, {lmax.xyz, rref}, {rmin.xyz, 0}, {rmax.xyz, 0}
// ref: internal child = node index; leaf = 0x80000000 | (count-1) << 27 | first slot.
// Stack entries are two words: (ref, entry distance of the deferred child).
__device__ __forceinline__ bool ref_is_leaf(uint32_t r) { return (r & 0x80000000u) != 0u; }

template <bool NEAREST, bool EARLY_EXIT, bool COUNT>
__device__ Hit trace_wide(const TraceScene &S, F3 ro, F3 rd, float t_max, LaneStack &stk, TraceCount &cnt) {
    Hit hit{kInf, 0u};
    uint32_t best_slot = 0xFFFFFFFFu;
    const F3 inv = fs(1.0f) / rd;
    if (S.n_nodes == 0u) return hit;
    float te, ta;
    slab(xyz(S.root_min), xyz(S.root_max), ro, inv, te, ta);
    if (COUNT) cnt.add(kCntNodes, 1);
    if (!(te <= ta && ta > 0.0f)) return hit;
    uint32_t sp = 0u;
    uint32_t cur = __float_as_uint(S.root_min.w);  // root ref (internal 0, or a leaf ref)
    for (;;) {
        if (COUNT) count_traversal_step(cnt);
        if (ref_is_leaf(cur)) {
            const uint32_t first = cur & 0x07FFFFFFu;
            const uint32_t count = ((cur >> 27) & 15u) + 1u;
            for (uint32_t i = 0u; i < count; i++) {
                const uint32_t slot = first + i;
                const float4 q0 = S.slot_tris[3 * slot], q1 = S.slot_tris[3 * slot + 1], q2 = S.slot_tris[3 * slot + 2];
                if (COUNT) cnt.add(kCntTris, 1);
                const float t = mt_slot(ro, rd, xyz(q0), xyz(q1), xyz(q2));
                if (!NEAREST) {
                    if (t < hit.t) {
                        hit.t = t;
                        hit.prim = __float_as_uint(q0.w);
                    }
                } else if (t < hit.t || (t == hit.t && t < kInf && slot < best_slot)) {
                    hit.t = t;
                    hit.prim = __float_as_uint(q0.w);
                    best_slot = slot;
                }
            }
            if (EARLY_EXIT && hit.t < t_max) return hit;
        } else {
            const float4 *w = S.wnodes + 4u * cur;
            const float4 a = w[0], b = w[1], c = w[2], d = w[3];
            if (COUNT) cnt.add(kCntNodes, 2);
            float tel, tal, ter, tar;
            slab(xyz(a), xyz(b), ro, inv, tel, tal);
            slab(xyz(c), xyz(d), ro, inv, ter, tar);
            const bool hl = tel <= tal && tal > 0.0f && (NEAREST ? tel <= hit.t : tel < hit.t);
            const bool hr = ter <= tar && tar > 0.0f && (NEAREST ? ter <= hit.t : ter < hit.t);
            const uint32_t lref = __float_as_uint(a.w), rref = __float_as_uint(b.w);
            if (hl && hr) {
                const bool lfirst = NEAREST ? (tel <= ter) : true;
                stk.put2(sp, lfirst ? rref : lref, __float_as_uint(lfirst ? ter : tel));
                sp += 2u;
                cur = lfirst ? lref : rref;
                continue;
            }
            if (hl || hr) {
                cur = hl ? lref : rref;
                continue;
            }
        }
        // pop, skipping deferred children whose entry is no longer before the best hit
        bool found = false;
        while (sp > 0u) {
            sp -= 2u;
            const uint2 e = stk.get2(sp);
            const float et = __uint_as_float(e.y);
            if (NEAREST ? et <= hit.t : et < hit.t) {
                cur = e.x;
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return hit;
}

// 4-wide traversal. Each record stands for a BVH2 node and holds up to four
// of its descendants (a cut of its subtree, left-to-right order kept), so the
// leaves are met in the reference's order. The BVH2 nodes between a record and
// its children are not tested. That changes no decision: every child box lies
// inside its parent's box (checked when the records are built), so per axis the
// slab interval of a child lies inside that of a skipped node, and a skipped
// node that would fail (te > ta, ta <= 0, or te >= best t at its visit) has a
// descendant child that fails the same way no later than the reference would
// test it. One exception comes from 1/rd = +-inf (a zero direction component):
// a box with a face in the ray's plane then gets a NaN slab end and fails,
// while a box flat on that axis and lying in the plane gets two NaN ends and
// ignores the axis. Each child carries one "kill" bit per axis, set at build
// time when it is flat on that axis at a face of a skipped non-flat node; such
// a child fails when 1/rd is infinite on that axis and ro lies in its plane,
// exactly when the skipped node would have failed.
// Deferred children keep their entry distance and are re-checked at pop time,
// as in trace_wide.
//
// Record (128 bytes): min.x[4], min.y[4], min.z[4], max.x[4], max.y[4],
// max.z[4] (SoA over the children), refs[4] (absent child = 0xFFFFFFFF),
// kill bits (3 per child: x, y, z), 12 bytes of padding.
//
// WT: the slot triangles hold the vertices (v0, v1, v2) and are tested with
// the watertight test instead of Moller-Trumbore, and boxes with the
// conservative slab_wt (traversal mode 4): the closest hit is the one a
// watertight test of every triangle finds.
//
// The traversal is written as a resumable query: Trav4 holds one query's
// state (ray, best hit, current record, stack depth; the stack itself lives in
// the lane's LaneStack), trav4_begin tests the root box, and a step is a
// leaf's triangles or a record's four children, then the pops up to the next
// entry to visit. trace_wide4 runs the steps to the end.
template <bool WT>
struct Trav4 {
    F3 ro, rd, inv;
    float lim;      // stop once the best hit is nearer than this (0: closest-hit query)
    Hit hit;
    uint32_t cur, sp;
    uint32_t inf4;  // kill-bit mask of the axes whose 1/rd is infinite, for the four children
    WtRay wr;       // WT only
    float slack;    // WT only
};

// Starts a query; returns true when it is already answered (the root box is
// missed, or the scene is empty). early_lim > 0 makes it an early-exit query.
template <bool WT, bool COUNT>
__device__ __forceinline__ bool trav4_begin(const TraceScene &S, Trav4<WT> &q, F3 ro, F3 rd, float early_lim,
                                            TraceCount &cnt) {
    q.ro = ro;
    q.rd = rd;
    q.inv = fs(1.0f) / rd;
    q.lim = early_lim;
    q.hit = Hit{kInf, 0u};
    q.sp = 0u;
    if (WT) q.wr = wt_ray(rd);
    if (S.n_nodes == 0u) return true;
    const uint32_t infmask = (__builtin_isinf(q.inv.x) ? 1u : 0u) | (__builtin_isinf(q.inv.y) ? 2u : 0u) |
                             (__builtin_isinf(q.inv.z) ? 4u : 0u);
    // the kill bits reproduce the reference's culling of rays in a box's min
    // face plane; the watertight mode's conservative slab keeps those boxes
    q.inf4 = WT ? 0u : infmask * 0x249u;  // replicated for the four children
    float te, ta;
    if (WT) {
        q.slack = wt_slack(ro, xyz(S.root_min), xyz(S.root_max));
        slab_wt(xyz(S.root_min), xyz(S.root_max), ro, q.inv, q.slack, te, ta);
    } else {
        slab(xyz(S.root_min), xyz(S.root_max), ro, q.inv, te, ta);
    }
    if (COUNT) cnt.add(kCntNodes, 1);
    q.cur = __float_as_uint(S.root_min.w);
    return !(te <= ta && ta > 0.0f);
}

// The parts of a traversal step of a started query. trav4_leaf tests the
// triangles of the leaf q.cur; it returns true when an early-exit query is answered.
// trav4_node tests the four children of record q.cur and returns true when it
// descended (q.cur set, the other passing children pushed). trav4_pop pops up
// to the next entry still nearer than the best hit; it returns true when the
// stack is empty (the query is answered).
template <bool WT, bool COUNT>
__device__ __forceinline__ bool trav4_leaf(const TraceScene &S, Trav4<WT> &q, TraceCount &cnt) {
    const uint32_t first = q.cur & 0x07FFFFFFu;
    const uint32_t count = ((q.cur >> 27) & 15u) + 1u;
    // two triangles per step: both loads in flight before either test;
    // the tests still run in slot order
    for (uint32_t i = 0u; i < count; i += 2u) {
        const uint32_t slot = first + i;
        const bool two = i + 1u < count;
        const float4 *p = S.slot_tris + 3u * slot;
        // the second triangle's loads are issued with the first's (from the
        // first slot again when there is no second: same lines, result unused)
        const float4 *pb = p + (two ? 3u : 0u);
        const float4 a0 = p[0], a1 = p[1], a2 = p[2];
        const float4 b0 = pb[0], b1 = pb[1], b2 = pb[2];
        if (COUNT) {
            cnt.add(kCntTris, 1);
            if (two) cnt.add(kCntTris, 1);
        }
        const float t = WT ? wt_slot(q.wr, q.ro, xyz(a0), xyz(a1), xyz(a2))
                           : mt_slot(q.ro, q.rd, xyz(a0), xyz(a1), xyz(a2));
        if (t < q.hit.t) {
            q.hit.t = t;
            q.hit.prim = __float_as_uint(a0.w);
        }
        if (two) {
            const float u = WT ? wt_slot(q.wr, q.ro, xyz(b0), xyz(b1), xyz(b2))
                               : mt_slot(q.ro, q.rd, xyz(b0), xyz(b1), xyz(b2));
            if (u < q.hit.t) {
                q.hit.t = u;
                q.hit.prim = __float_as_uint(b0.w);
            }
        }
    }
    return q.hit.t < q.lim;
}

template <bool WT, bool COUNT>
__device__ __forceinline__ bool trav4_node(const TraceScene &S, Trav4<WT> &q, LaneStack &stk, TraceCount &cnt) {
    const float4 *w = S.w4nodes + 8u * q.cur;
    const float4 mnx = w[0], mny = w[1], mnz = w[2], mxx = w[3], mxy = w[4], mxz = w[5];
    const uint4 refs = *reinterpret_cast<const uint4 *>(&w[6]);
    const uint32_t kill = q.inf4 != 0u ? __float_as_uint(w[7].x) & q.inf4 : 0u;
    const float bmnx[4] = {mnx.x, mnx.y, mnx.z, mnx.w}, bmny[4] = {mny.x, mny.y, mny.z, mny.w};
    const float bmnz[4] = {mnz.x, mnz.y, mnz.z, mnz.w}, bmxx[4] = {mxx.x, mxx.y, mxx.z, mxx.w};
    const float bmxy[4] = {mxy.x, mxy.y, mxy.z, mxy.w}, bmxz[4] = {mxz.x, mxz.y, mxz.z, mxz.w};
    const uint32_t r[4] = {refs.x, refs.y, refs.z, refs.w};
    float tes[4];
    uint32_t pass = 0u;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float e, a;
        if (WT) slab_wt(f3(bmnx[k], bmny[k], bmnz[k]), f3(bmxx[k], bmxy[k], bmxz[k]), q.ro, q.inv, q.slack, e, a);
        else slab(f3(bmnx[k], bmny[k], bmnz[k]), f3(bmxx[k], bmxy[k], bmxz[k]), q.ro, q.inv, e, a);
        tes[k] = e;
        // every condition evaluated (no short-circuit branches)
        const bool ok = (r[k] != 0xFFFFFFFFu) & (e <= a) & (a > 0.0f) & (e < q.hit.t);
        pass |= (uint32_t)ok << k;
    }
    // kill bits only matter for rays with a zero direction component
    // (rare): one branch per record instead of one per child
    if (kill != 0u) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t kb = (kill >> (3 * k)) & 7u;
            if (((kb & 1u) && q.ro.x == bmnx[k]) || ((kb & 2u) && q.ro.y == bmny[k]) ||
                ((kb & 4u) && q.ro.z == bmnz[k]))
                pass &= ~(1u << k);
        }
    }
    if (COUNT) cnt.add(kCntNodes, 4);
    if (pass != 0u) {
        // push passing children after the first, rightmost first; when no lane
        // can reach the spill words (at most three entries are pushed) the
        // pushes go straight to LDS
        const bool shallow = q.sp + 6u <= (uint32_t)kLdsStack;
        if (__ballot(!shallow) == 0ull) {
            // every candidate is written at the top, which only advances
            // past the pushed ones (the slots written stay below kLdsStack)
#pragma unroll
            for (int k = 3; k >= 1; k--) {
                stk.put2_lds(q.sp, r[k], __float_as_uint(tes[k]));
                q.sp += (((pass >> k) & 1u) && (pass & ((1u << k) - 1u))) ? 2u : 0u;
            }
        } else {
#pragma unroll
            for (int k = 3; k >= 1; k--) {
                if (((pass >> k) & 1u) && (pass & ((1u << k) - 1u))) {
                    stk.put2(q.sp, r[k], __float_as_uint(tes[k]));
                    q.sp += 2u;
                }
            }
        }
        const int f = __builtin_ctz(pass);
        q.cur = f == 0 ? r[0] : f == 1 ? r[1] : f == 2 ? r[2] : r[3];
        return true;
    }
    return false;
}

template <bool WT>
__device__ __forceinline__ bool trav4_pop(Trav4<WT> &q, LaneStack &stk) {
    if (__ballot(q.sp > (uint32_t)kLdsStack) == 0ull) {
        // no lane's entries reach the spill words
        while (q.sp > 0u) {
            q.sp -= 2u;
            const uint2 e = stk.get2_lds(q.sp);
            if (__uint_as_float(e.y) < q.hit.t) {
                q.cur = e.x;
                return false;
            }
        }
        return true;
    }
    while (q.sp > 0u) {
        q.sp -= 2u;
        const uint2 e = stk.get2(q.sp);
        if (__uint_as_float(e.y) < q.hit.t) {
            q.cur = e.x;
            return false;
        }
    }
    return true;
}

// One step of a started query (a leaf or a record, then the pops); returns
// true when the query is answered. Lanes of a wavefront at a leaf and lanes at
// a record run their steps in the same loop iteration.
template <bool WT, bool COUNT>
__device__ __forceinline__ bool trav4_step(const TraceScene &S, Trav4<WT> &q, LaneStack &stk, TraceCount &cnt) {
    if (COUNT) count_traversal_step(cnt);
    if (ref_is_leaf(q.cur)) {
        if (trav4_leaf<WT, COUNT>(S, q, cnt)) return true;
    } else if (trav4_node<WT, COUNT>(S, q, stk, cnt)) {
        return false;
    }
    return trav4_pop(q, stk);
}

template <bool EARLY_EXIT, bool COUNT, bool WT = false>
__device__ Hit trace_wide4(const TraceScene &S, F3 ro, F3 rd, float t_max, LaneStack &stk, TraceCount &cnt) {
    Trav4<WT> q;
    if (!trav4_begin<WT, COUNT>(S, q, ro, rd, EARLY_EXIT ? t_max : 0.0f, cnt))
        while (!trav4_step<WT, COUNT>(S, q, stk, cnt)) {
        }
    return q.hit;
}

// Traversal modes: 0 = the reference's BVH2 nodes, 1 = child-box records in the
// reference's order, 2 = child-box records nearest child first, 3 = 4-wide
// records in the reference's order (default), 4 = 4-wide records with the
// watertight triangle test (slot triangles as vertices).
template <int MODE, bool EARLY_EXIT, bool COUNT>
__device__ __forceinline__ Hit trace_mode(const TraceScene &S, F3 ro, F3 rd, float t_max, LaneStack &stk,
                                          TraceCount &cnt) {
    if (MODE == 0) return trace_reference<EARLY_EXIT, COUNT>(S, ro, rd, t_max, stk, cnt);
    if (MODE == 3) return trace_wide4<EARLY_EXIT, COUNT>(S, ro, rd, t_max, stk, cnt);
    if (MODE == 4) return trace_wide4<EARLY_EXIT, COUNT, true>(S, ro, rd, t_max, stk, cnt);
    return trace_wide<MODE == 2, EARLY_EXIT, COUNT>(S, ro, rd, t_max, stk, cnt);
}

}  // namespace rtd
