/*
 * rt_types.h — byte images of the records rustracer uploads to the GPU.
 *
 * Every struct here is the C mirror of a `#[repr(C)]` Rust type of the
 * reference and of the WGSL struct that reads it. Sizes and field offsets are
 * the upload contract, so they are pinned with static assertions below:
 *   rt_triangle        <- TriangleGpu    rustracer-core/src/scene.rs:74-94        (128 B)
 *   rt_bvh_node        <- BVHNode        rustracer-core/src/bvh.rs:79-88          (48 B)
 *   rt_material        <- MaterialGpu    rustracer-core/src/material.rs:44-58     (64 B)
 *   rt_light           <- LightGpu       rustracer-core/src/scene.rs:141-150      (64 B)
 *   rt_camera_uniform  <- CameraUniform  rustracer-core/src/camera.rs:35-52       (96 B)
 *   rt_photon          <- PhotonGpu      rustracer-renderer/src/lib.rs:86-97      (64 B)
 *   rt_image_params    <- ImageParams    rustracer-renderer/src/lib.rs:67-84      (80 B)
 *   rt_grid_params     <- GridParamsUniform lib.rs:137-150                        (48 B)
 *   rt_denoise_params  <- DenoiseParamsUniform lib.rs:113-124                     (32 B)
 *   rt_composite_params<- CompositeParams lib.rs:126-133                          (16 B)
 *   rt_tonemap_uniform <- ToneMapUniform lib.rs:99-111                            (32 B)
 * Sizes are the ones rustracer-renderer/tests/gpu_layout.rs:12-41 checks;
 * MaterialGpu offsets are the ones rustracer-core/src/material.rs:193-216 checks.
 *
 * Plain C (no HIP or torch types) so that any FFI can bind it.
 */
#ifndef RT_TYPES_H
#define RT_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* MaterialKind discriminants, rustracer-core/src/material.rs:13-20 */
#define RT_MAT_LAMBERTIAN 0u
#define RT_MAT_METAL 1u
#define RT_MAT_DIELECTRIC 2u
#define RT_MAT_EMISSIVE 3u
/* TEX_NONE = u32::MAX, material.rs:7; MAT_FLAG_ALPHA_CUTOUT, material.rs:10 */
#define RT_TEX_NONE 0xFFFFFFFFu
#define RT_MAT_FLAG_ALPHA_CUTOUT 1u

/* Texture roles (lib.rs:502-507) and one role's rgba8 array (device or host). */
#define RT_TEX_ROLE_ALBEDO 0u
#define RT_TEX_ROLE_MR 1u
#define RT_TEX_ROLE_EMISSIVE 2u
#define RT_TEX_ROLE_NORMAL 3u
#define RT_TEX_ROLES 4u
typedef struct rt_tex_array {
    const uint8_t *texels; /* layers * height * width * 4 bytes, row-major */
    uint32_t width, height, layers, _pad;
} rt_tex_array;

/* LightGpu.kind values, scene.rs:147 and lib.rs:1231-1282 */
#define RT_LIGHT_POINT 0u
#define RT_LIGHT_DIRECTIONAL 1u
#define RT_LIGHT_AREA 2u
#define RT_LIGHT_ENVIRONMENT 3u

/* Photon grid limits of the reference, lib.rs:15-19 */
#define RT_GRID_MAX_DIM 26u
#define RT_GRID_MAX_CELLS (26u * 26u * 26u)
#define RT_MAX_BOUNCES 6u

typedef struct rt_triangle {
    float v0[3];
    float _pad0;
    float v1[3];
    float _pad1;
    float v2[3];
    float _pad2;
    float n0[3];
    float _pad3;
    float n1[3];
    float _pad4;
    float n2[3];
    float _pad5;
    float uv0[2];
    float uv1[2];
    float uv2[2];
    uint32_t material_id;
    float _pad6;
} rt_triangle;

typedef struct rt_bvh_node {
    float bbox_min[3];
    uint32_t left_first; /* left child (internal) or first prim slot (leaf) */
    float bbox_max[3];
    uint32_t right_child; /* right child (internal), unused for leaves */
    uint32_t prim_count;  /* 0 = internal node */
    float _pad[3];
} rt_bvh_node;

typedef struct rt_material {
    float albedo[4];
    float emissive[3];
    float roughness;
    float metallic;
    float ior;
    uint32_t kind;
    uint32_t flags;
    uint32_t tex_albedo;
    uint32_t tex_mr;
    uint32_t tex_emissive;
    uint32_t tex_normal;
} rt_material;

typedef struct rt_light {
    float data[4]; /* point: position; directional: direction; area: triangle index as f32;
                      environment: horizon colour */
    float color[3]; /* environment: zenith colour */
    float intensity;
    uint32_t kind;
    float _pad_a[3];
    float _pad_b[4];
} rt_light;

typedef struct rt_camera_uniform {
    float position[3];
    float _pad0;
    float forward[3];
    float _pad1;
    float right[3]; /* pre-scaled by aspect * tan(fov/2) */
    float _pad2;
    float up[3]; /* pre-scaled by tan(fov/2) */
    float _pad3;
    float inv_half_res[2];
    float near_plane;
    float fov_scale;
    float aperture;
    float focus_dist;
    float _pad_end[2];
} rt_camera_uniform;

typedef struct rt_photon {
    float position[3];
    float _pad0;
    float power[3];
    float _pad1;
    float incident[3];
    float _pad2;
    float normal[3];
    float _pad3;
} rt_photon;

typedef struct rt_image_params {
    uint32_t width;
    uint32_t height;
    uint32_t spp;
    uint32_t frame;
    float light_emission[3];
    float ambient;
    float accumulate_alpha;
    uint32_t use_photons;
    uint32_t photon_count;
    float photon_radius;
    float photon_scale;
    uint32_t photon_debug;
    float point_light_emission[3];
    uint32_t _pad[3]; /* _pad[0] carries debug_view, lib.rs:345 */
} rt_image_params;

typedef struct rt_grid_params {
    float room_min_x;
    float room_min_y;
    float room_min_z;
    float cell_size;
    uint32_t grid_x;
    uint32_t grid_y;
    uint32_t grid_z;
    uint32_t num_cells;
    uint32_t photon_slots;
    uint32_t _pad[3];
} rt_grid_params;

typedef struct rt_denoise_params {
    uint32_t width;
    uint32_t height;
    uint32_t step_size;
    float lambda;
    float phi_normal;
    float phi_depth;
    float alpha;
    uint32_t _pad;
} rt_denoise_params;

typedef struct rt_composite_params {
    uint32_t width;
    uint32_t height;
    float alpha;
    uint32_t _pad;
} rt_composite_params;

typedef struct rt_tonemap_uniform {
    float exposure;
    float gamma;
    uint32_t width;
    float glow;
    float marker[4];
} rt_tonemap_uniform;

/* Closest-hit result of one ray query: t >= 1e38 means miss (gather.wgsl:66-88). */
typedef struct rt_hit {
    float t;
    uint32_t prim_id;
} rt_hit;

/* One ray of a fixed ray set: origin then direction. */
typedef struct rt_ray {
    float origin[3];
    float dir[3];
} rt_ray;

#ifdef __cplusplus
} /* extern "C" */
#define RT_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define RT_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

RT_STATIC_ASSERT(sizeof(rt_triangle) == 128, "TriangleGpu is 128 bytes");
RT_STATIC_ASSERT(sizeof(rt_bvh_node) == 48, "BVHNode is 48 bytes");
RT_STATIC_ASSERT(sizeof(rt_material) == 64, "MaterialGpu is 64 bytes");
RT_STATIC_ASSERT(sizeof(rt_light) == 64, "LightGpu is 64 bytes");
RT_STATIC_ASSERT(sizeof(rt_camera_uniform) == 96, "CameraUniform is 96 bytes");
RT_STATIC_ASSERT(sizeof(rt_photon) == 64, "PhotonGpu is 64 bytes");
RT_STATIC_ASSERT(sizeof(rt_image_params) == 80, "ImageParams is 80 bytes");
RT_STATIC_ASSERT(sizeof(rt_grid_params) == 48, "GridParamsUniform is 48 bytes");
RT_STATIC_ASSERT(sizeof(rt_denoise_params) == 32, "DenoiseParamsUniform is 32 bytes");
RT_STATIC_ASSERT(sizeof(rt_composite_params) == 16, "CompositeParams is 16 bytes");
RT_STATIC_ASSERT(sizeof(rt_tonemap_uniform) == 32, "ToneMapUniform is 32 bytes");
RT_STATIC_ASSERT(offsetof(rt_material, roughness) == 28, "material.rs:209");
RT_STATIC_ASSERT(offsetof(rt_material, metallic) == 32, "material.rs:210");
RT_STATIC_ASSERT(offsetof(rt_material, kind) == 40, "material.rs:211");
RT_STATIC_ASSERT(offsetof(rt_material, flags) == 44, "material.rs:212");
RT_STATIC_ASSERT(offsetof(rt_material, tex_albedo) == 48, "material.rs:213");
RT_STATIC_ASSERT(offsetof(rt_material, tex_emissive) == 56, "material.rs:214");
RT_STATIC_ASSERT(offsetof(rt_material, tex_normal) == 60, "material.rs:215");
RT_STATIC_ASSERT(offsetof(rt_light, kind) == 32, "scene.rs:128-140 layout comment");
RT_STATIC_ASSERT(offsetof(rt_triangle, material_id) == 120, "scene.rs:92");
RT_STATIC_ASSERT(offsetof(rt_bvh_node, prim_count) == 32, "bvh.rs:86");
RT_STATIC_ASSERT(offsetof(rt_camera_uniform, inv_half_res) == 64, "camera.rs:47");
RT_STATIC_ASSERT(offsetof(rt_image_params, accumulate_alpha) == 32, "lib.rs:76");
RT_STATIC_ASSERT(offsetof(rt_image_params, point_light_emission) == 56, "lib.rs:82");
RT_STATIC_ASSERT(offsetof(rt_tonemap_uniform, marker) == 16, "lib.rs:110");

#endif /* RT_TYPES_H */
