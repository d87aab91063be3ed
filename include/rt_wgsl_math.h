/*
 * rt_wgsl_math.h — one fixed definition of the WGSL builtins whose results
 * the WGSL specification leaves to the implementation.
 *
 * The reference's shaders call sin/cos (common.wgsl:60-67, gather.wgsl:247-249),
 * pow (common.wgsl:24, gather.wgsl:110,423, denoise.wgsl:44, tonemap.wgsl:178),
 * exp (denoise.wgsl:48,110), inverseSqrt (gather.wgsl:244), smoothstep
 * (photon_trace.wgsl:93) and the float->int conversions (photon_count.wgsl:18-23,
 * gather.wgsl:55-58,155). WGSL fixes only error bounds for the transcendentals
 * (absolute 2^-11 for sin/cos on [-pi, pi]; exp/exp2 3+2|x| ULP; pow inherited
 * from exp2(y*log2(x))), so every GPU driver answers differently in the last
 * bits. This backend picks one answer and writes it down here, built only from
 * IEEE-754 +, -, *, /, fma, sqrt, rint/floor and bit operations, so that the HIP
 * kernels (compiled with -ffp-contract=off) and a CPU build of the same
 * functions produce identical bits.
 *
 * Polynomials are evaluated by Horner's rule with fused multiply-adds
 * (fmaf: one rounding per step on the host, where C's fmaf is the IEEE
 * operation, and on gfx950, v_fma_f32), the rest with single operations.
 *
 * Methods (own implementations of textbook algorithms):
 *   sin/cos : Cody-Waite reduction by pi/2 with a 3-part split of pi/2
 *             (8 + 11 significant bits, then the rounded rest), then the Taylor
 *             polynomials of sin (to r^9) and cos (to r^10) on |r| <= pi/4.
 *   exp2    : x = n + f, |f| <= 1/2, 2^f by the degree-7 Taylor polynomial of
 *             e^(f ln 2), scaled by 2^n through the exponent field.
 *   log2    : x = m 2^e with m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s),
 *             s = (m-1)/(m+1), by the odd series to s^11.
 *   exp     : exp2(x * log2(e)).   pow: exp2(y * log2(x)).
 * Accuracy (checked against double-precision libm by the test suite): sin/cos
 * absolute error < 1e-6 on [-8192, 8192]; exp2/log2 relative error < 4e-7.
 */
#ifndef RT_WGSL_MATH_H
#define RT_WGSL_MATH_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define RT_WM_FN __host__ __device__ static inline
#else
#define RT_WM_FN static inline
#endif

RT_WM_FN uint32_t wm_f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
RT_WM_FN float wm_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
RT_WM_FN float wm_u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

/* min/max with one NaN rule: a NaN operand yields the other operand; for equal
 * operands (including -0 vs +0) the first operand is returned. */
RT_WM_FN float wmin(float a, float b) { return (b < a) ? b : ((a == a) ? a : b); }
RT_WM_FN float wmax(float a, float b) { return (b > a) ? b : ((a == a) ? a : b); }
RT_WM_FN float wclamp(float x, float lo, float hi) { return wmin(wmax(x, lo), hi); }

/* WGSL u32(f32) / i32(f32): truncation toward zero, saturating, NaN -> 0. */
RT_WM_FN uint32_t wgsl_u32(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
RT_WM_FN int32_t wgsl_i32(float x) {
    if (x != x) return 0;
    if (x <= -2147483648.0f) return (int32_t)0x80000000u;
    if (x >= 2147483648.0f) return 0x7FFFFFFF;
    return (int32_t)x;
}

/* Correctly rounded square root (IEEE-754 sqrt on both host and gfx950). */
RT_WM_FN float wgsl_sqrt(float x) { return __builtin_sqrtf(x); }
RT_WM_FN float wgsl_inverse_sqrt(float x) { return 1.0f / __builtin_sqrtf(x); }

RT_WM_FN float wm_sin_poly(float r) {
    const float r2 = r * r;
    float p = 2.75573192e-06f;                   /*  1/9! */
    p = wm_fma(p, r2, -1.98412698e-04f);         /* -1/7! */
    p = wm_fma(p, r2, 8.33333377e-03f);          /*  1/5! */
    p = wm_fma(p, r2, -1.66666672e-01f);         /* -1/3! */
    return wm_fma(r * r2, p, r);
}
RT_WM_FN float wm_cos_poly(float r) {
    const float r2 = r * r;
    float p = -2.75573188e-07f;                  /* -1/10! */
    p = wm_fma(p, r2, 2.48015876e-05f);          /*  1/8!  */
    p = wm_fma(p, r2, -1.38888892e-03f);         /* -1/6!  */
    p = wm_fma(p, r2, 4.16666679e-02f);          /*  1/4!  */
    return wm_fma(r2 * r2, p, 1.0f - 0.5f * r2);
}

/* Reduce x to r in [-pi/4, pi/4] and the quadrant k mod 4. Valid for |x| <= 8192. */
RT_WM_FN float wm_reduce_half_pi(float x, uint32_t *quadrant) {
    const float k = __builtin_rintf(x * 0.636619747f); /* 2/pi */
    float r = x - k * 1.5703125f;                      /* 8 significant bits: exact */
    r = r - k * 4.83751297e-04f;                       /* 11 significant bits: exact */
    r = r - k * 7.54979013e-08f;                       /* rest of pi/2, rounded */
    *quadrant = (uint32_t)(int32_t)k & 3u;
    return r;
}

RT_WM_FN float wgsl_sin(float x) {
    if (!(__builtin_fabsf(x) <= 8192.0f)) return __builtin_nanf("");
    uint32_t q;
    const float r = wm_reduce_half_pi(x, &q);
    const float s = wm_sin_poly(r), c = wm_cos_poly(r);
    return q == 0u ? s : q == 1u ? c : q == 2u ? -s : -c;
}
RT_WM_FN float wgsl_cos(float x) {
    if (!(__builtin_fabsf(x) <= 8192.0f)) return __builtin_nanf("");
    uint32_t q;
    const float r = wm_reduce_half_pi(x, &q);
    const float s = wm_sin_poly(r), c = wm_cos_poly(r);
    return q == 0u ? c : q == 1u ? -s : q == 2u ? -c : s;
}

RT_WM_FN float wgsl_exp2(float x) {
    if (x != x) return x;
    if (x >= 128.0f) return __builtin_inff();
    if (x < -151.0f) return 0.0f;
    const float n = __builtin_rintf(x);
    const float f = x - n; /* exact, |f| <= 1/2 */
    float p = 1.52527336e-05f;                   /* ln2^7/7! */
    p = wm_fma(p, f, 1.54035297e-04f);           /* ln2^6/6! */
    p = wm_fma(p, f, 1.33335579e-03f);           /* ln2^5/5! */
    p = wm_fma(p, f, 9.61812865e-03f);           /* ln2^4/4! */
    p = wm_fma(p, f, 5.55041097e-02f);           /* ln2^3/3! */
    p = wm_fma(p, f, 2.40226507e-01f);           /* ln2^2/2! */
    p = wm_fma(p, f, 6.93147182e-01f);           /* ln2      */
    p = wm_fma(p, f, 1.0f);
    const int32_t e = (int32_t)n;
    if (e >= -126) {
        if (e > 127) return (p * 1.70141183e+38f) * 2.0f; /* 2^127 * 2 */
        return p * wm_u2f((uint32_t)(e + 127) << 23);
    }
    /* subnormal result: one exact scaling into the normal range, one rounding */
    return (p * wm_u2f((uint32_t)(e + 64 + 127) << 23)) * 5.42101086e-20f; /* 2^-64 */
}

RT_WM_FN float wgsl_log2(float x) {
    if (x != x) return x;
    if (x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_inff();
    if (x == __builtin_inff()) return x;
    int32_t bias = 0;
    if (x < 1.17549435e-38f) { x = x * 8388608.0f; bias = -23; } /* subnormal input */
    const uint32_t u = wm_f2u(x);
    int32_t e = (int32_t)((u >> 23) & 0xFFu) - 127 + bias;
    float m = wm_u2f((u & 0x007FFFFFu) | 0x3F800000u); /* [1, 2) */
    if (m > 1.41421354f) { m = m * 0.5f; e = e + 1; }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p = 9.09090936e-02f;             /* 1/11 */
    p = wm_fma(p, s2, 1.11111112e-01f);    /* 1/9 */
    p = wm_fma(p, s2, 1.42857149e-01f);    /* 1/7 */
    p = wm_fma(p, s2, 2.00000003e-01f);    /* 1/5 */
    p = wm_fma(p, s2, 3.33333343e-01f);    /* 1/3 */
    const float ln_m = wm_fma(2.0f * s, s2 * p, 2.0f * s);
    return wm_fma(ln_m, 1.44269502f, (float)e); /* log2(e) */
}

RT_WM_FN float wgsl_exp(float x) { return wgsl_exp2(x * 1.44269502f); }
RT_WM_FN float wgsl_pow(float x, float y) { return wgsl_exp2(y * wgsl_log2(x)); }

RT_WM_FN float wgsl_smoothstep(float e0, float e1, float x) {
    const float t = wclamp((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

#endif /* RT_WGSL_MATH_H */
