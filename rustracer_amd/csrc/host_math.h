// host_math.h — host-side vector math with glam's evaluation order
// (dot = (x*x' + y*y') + z*z', normalize = v * (1 / length)), used by the
// scene model, BVH builder and camera so that host-built records match the
// values rustracer-core computes with glam (FMA contraction off).
#pragma once

#include <cmath>
#include <cstdint>

namespace rt {

struct Vec3 {
    float x = 0.f, y = 0.f, z = 0.f;
    Vec3() = default;
    constexpr Vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
struct Vec2 {
    float x = 0.f, y = 0.f;
    Vec2() = default;
    constexpr Vec2(float a, float b) : x(a), y(b) {}
};

inline Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator*(Vec3 a, Vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline Vec3 operator*(Vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec3 operator*(float s, Vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline Vec3 operator/(Vec3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline Vec3 operator-(Vec3 a) { return {-a.x, -a.y, -a.z}; }
inline Vec3 &operator+=(Vec3 &a, Vec3 b) { a = a + b; return a; }
inline float dot(Vec3 a, Vec3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline float length(Vec3 v) { return std::sqrt(dot(v, v)); }
inline Vec3 normalize(Vec3 v) { return v * (1.0f / length(v)); }
inline Vec3 normalize_or_zero(Vec3 v) {
    const float r = 1.0f / length(v);
    if (std::isfinite(r) && r > 0.0f) return v * r;
    return {0.f, 0.f, 0.f};
}
inline Vec3 vmin(Vec3 a, Vec3 b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
inline Vec3 vmax(Vec3 a, Vec3 b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
inline float max_element(Vec3 v) { return std::fmax(std::fmax(v.x, v.y), v.z); }

// Rust's f32::to_radians: self * (PI / 180) evaluated in f32.
inline float to_radians(float deg) { return deg * (3.14159265358979323846f / 180.0f); }

// Rotation of `v` by the unit quaternion for (axis, angle), in the order
// glam's Quat::from_axis_angle and Quat * Vec3 use.
inline Vec3 rotate_axis_angle(Vec3 axis, float angle, Vec3 v) {
    const float s = std::sin(angle * 0.5f), c = std::cos(angle * 0.5f);
    const Vec3 b = axis * s;
    const float w = c;
    const float b2 = dot(b, b);
    return v * (w * w - b2) + b * (dot(v, b) * 2.0f) + cross(b, v) * (w * 2.0f);
}

// Seeded host RNG (SplitMix64) for the procedural scenes.
struct HostRng {
    uint64_t s;
    explicit HostRng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in [0, 1)
    float uniform() { return (float)((next() >> 40) * (1.0 / 16777216.0)); }
    float uniform(float lo, float hi) { return lo + (hi - lo) * uniform(); }
    float log_uniform(float lo, float hi) {
        return std::exp(std::log(lo) + (std::log(hi) - std::log(lo)) * uniform());
    }
    float normal() {
        const double u1 = ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        return (float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    }
    uint32_t below(uint32_t n) { return (uint32_t)(next() % n); }
};

}  // namespace rt
