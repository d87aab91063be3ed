// scene.h — host scene model (rustracer-core/src/scene.rs, bvh.rs, material.rs, camera.rs).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/rustracer_hip.h"
#include "host_math.h"

namespace rt {

// rustracer-core/src/scene.rs:15-21
struct Mesh {
    std::vector<Vec3> positions;
    std::vector<Vec3> normals;
    std::vector<Vec2> texcoords;
    std::vector<uint32_t> indices;
    uint32_t material_id = 0;
    size_t triangle_count() const { return indices.size() / 3; }
};

// rustracer-core/src/bvh.rs:8-72
struct AABB {
    Vec3 min{3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
    Vec3 max{-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
    void extend(Vec3 p) {
        min = vmin(min, p);
        max = vmax(max, p);
    }
    AABB unite(const AABB &o) const {
        AABB r;
        r.min = vmin(min, o.min);
        r.max = vmax(max, o.max);
        return r;
    }
    float surface_area() const {
        const Vec3 e = max - min;
        return 2.0f * (e.x * e.y + e.y * e.z + e.z * e.x);
    }
};

struct BVH {
    std::vector<rt_bvh_node> nodes;
    std::vector<uint32_t> prim_indices;
};

// rustracer-core/src/bvh.rs:120-336
BVH build_bvh(const std::vector<Vec3> &centroids, const std::vector<AABB> &boxes);
// Depth of the deepest leaf (root = 1), and whether every node is reachable once.
uint32_t bvh_depth(const BVH &bvh);

// rustracer-core/src/texture.rs:5-9: RGBA linear floats.
struct Texture {
    uint32_t width = 0, height = 0;
    std::vector<float> rgba;  // width * height * 4
};

// One role's texture array as lib.rs:1286-1358 uploads it: every layer has the
// size of the first texture, a texture fills its top-left corner, texels are
// (u8)(clamp(c, 0, 1) * 255); an empty role is one 1x1 zero layer.
struct TextureArray {
    uint32_t width = 1, height = 1, layers = 1;
    std::vector<uint8_t> texels = std::vector<uint8_t>(4, 0);
};
TextureArray build_texture_array(const std::vector<Texture> &textures);

struct Scene {
    std::vector<Mesh> meshes;
    std::vector<rt_material> materials;
    std::vector<rt_light_desc> lights;
    std::vector<Texture> textures[4];  // albedo, metallic-roughness, emissive, normal

    size_t triangle_count() const;
    bool bounds(Vec3 &bmin, Vec3 &bmax) const;
    std::vector<rt_triangle> flatten_triangles() const;
    BVH build_bvh() const;
    std::vector<rt_light> lights_gpu() const;
};

rt_light light_to_gpu(const rt_light_desc &l);

// rustracer-core/src/material.rs:101-161
rt_material material_lambertian(Vec3 albedo);
rt_material material_metal(Vec3 albedo, float roughness);
rt_material material_dielectric(float ior, Vec3 albedo);
rt_material material_emissive(Vec3 color, float intensity);

// rustracer-core/src/scene.rs:463-505
void add_quad(Scene &s, Vec3 v0, Vec3 v1, Vec3 v2, Vec3 v3, uint32_t material);
void add_cube(Scene &s, Vec3 lo, Vec3 hi, uint32_t material);

// rustracer-core/src/config.rs:62-174: parse a scene.toml descriptor and apply
// it (environment override, extra lights, material overrides). Throws
// ApiError(RT_ERR_INVALID) on a malformed file, leaving the scene unchanged.
void apply_descriptor(Scene &scene, const std::string &toml);

std::unique_ptr<Scene> builtin_scene(const std::string &name, uint64_t seed);
bool builtin_camera(const std::string &name, uint64_t seed, uint32_t index, rt_camera &out);

// camera.rs
rt_camera camera_default();
rt_camera_uniform camera_to_uniform(const rt_camera &c, uint32_t w, uint32_t h);
rt_camera camera_frame_scene(const Scene &s);
void camera_orbit(rt_camera &c, float yaw, float pitch);
void camera_zoom(rt_camera &c, float delta);
void camera_pan(rt_camera &c, float dx, float dy);
// lib.rs:865-885 Renderer::project_marker: pixel centre, pixel radius and camera
// distance (0 = hidden) of the light gizmo at `pos` for a w x h image.
void project_marker(const rt_camera &c, uint32_t w, uint32_t h, Vec3 pos, float radius, float out[4]);

}  // namespace rt
