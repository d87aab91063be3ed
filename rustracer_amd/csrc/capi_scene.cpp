// capi_scene.cpp — C ABI over the host scene model, BVH builder and camera
// (rustracer-core), plus the error channel. Host-only: no HIP calls.
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "capi_util.h"
#include "scene.h"

struct rt_scene {
    rt::Scene scene;
};

namespace rt {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }
const Scene &scene_of(const rt_scene *s) { return s->scene; }
Scene &scene_of(rt_scene *s) { return s->scene; }

}  // namespace rt

using rt::api_call;
using rt::ApiError;

extern "C" {

const char *rt_last_error(void) { return rt::g_last_error.c_str(); }
const char *rt_version(void) { return "rustracer-hip 0.1 (gfx950)"; }

int rt_scene_new(rt_scene **out) {
    return api_call([&] {
        if (!out) throw ApiError(RT_ERR_INVALID, "null out");
        *out = new rt_scene();
    });
}

void rt_scene_free(rt_scene *s) { delete s; }

int rt_scene_add_material(rt_scene *s, const rt_material *m, uint32_t *index) {
    return api_call([&] {
        if (!s || !m) throw ApiError(RT_ERR_INVALID, "null argument");
        s->scene.materials.push_back(*m);
        if (index) *index = (uint32_t)(s->scene.materials.size() - 1);
    });
}

int rt_scene_add_mesh(rt_scene *s, const float *positions, uint32_t n_positions, const float *normals,
                      uint32_t n_normals, const float *uvs, uint32_t n_uvs, const uint32_t *indices, uint32_t n_indices,
                      uint32_t material_id) {
    return api_call([&] {
        if (!s || (n_positions && !positions) || (n_normals && !normals) || (n_uvs && !uvs) ||
            (n_indices && !indices))
            throw ApiError(RT_ERR_INVALID, "null array with non-zero count");
        if (n_indices % 3 != 0) throw ApiError(RT_ERR_INVALID, "index count must be a multiple of 3");
        for (uint32_t i = 0; i < n_indices; i++)
            if (indices[i] >= n_positions) throw ApiError(RT_ERR_INVALID, "index out of range of positions");
        rt::Mesh m;
        m.positions.resize(n_positions);
        for (uint32_t i = 0; i < n_positions; i++)
            m.positions[i] = rt::Vec3{positions[3 * i], positions[3 * i + 1], positions[3 * i + 2]};
        m.normals.resize(n_normals);
        for (uint32_t i = 0; i < n_normals; i++)
            m.normals[i] = rt::Vec3{normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]};
        m.texcoords.resize(n_uvs);
        for (uint32_t i = 0; i < n_uvs; i++) m.texcoords[i] = rt::Vec2{uvs[2 * i], uvs[2 * i + 1]};
        m.indices.assign(indices, indices + n_indices);
        m.material_id = material_id;
        s->scene.meshes.push_back(std::move(m));
    });
}

int rt_scene_add_light(rt_scene *s, const rt_light_desc *l) {
    return api_call([&] {
        if (!s || !l) throw ApiError(RT_ERR_INVALID, "null argument");
        if (l->kind > RT_LIGHT_ENVIRONMENT) throw ApiError(RT_ERR_INVALID, "unknown light kind");
        s->scene.lights.push_back(*l);
    });
}

int rt_scene_clear_lights(rt_scene *s) {
    return api_call([&] {
        if (!s) throw ApiError(RT_ERR_INVALID, "null scene");
        s->scene.lights.clear();
    });
}

int rt_scene_builtin(const char *name, uint64_t seed, rt_scene **out) {
    return api_call([&] {
        if (!name || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        auto sc = rt::builtin_scene(name, seed);
        if (!sc) throw ApiError(RT_ERR_INVALID, std::string("unknown built-in scene: ") + name);
        auto *w = new rt_scene();
        w->scene = std::move(*sc);
        *out = w;
    });
}

int rt_scene_builtin_camera(const char *name, uint64_t seed, uint32_t index, rt_camera *out) {
    return api_call([&] {
        if (!name || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        if (!rt::builtin_camera(name, seed, index, *out))
            throw ApiError(RT_ERR_INVALID, std::string("unknown built-in scene: ") + name);
    });
}

int rt_scene_counts(const rt_scene *s, uint32_t *nt, uint32_t *nm, uint32_t *nl, uint32_t *nmesh) {
    return api_call([&] {
        if (!s) throw ApiError(RT_ERR_INVALID, "null scene");
        if (nt) *nt = (uint32_t)s->scene.triangle_count();
        if (nm) *nm = (uint32_t)s->scene.materials.size();
        if (nl) *nl = (uint32_t)s->scene.lights.size();
        if (nmesh) *nmesh = (uint32_t)s->scene.meshes.size();
    });
}

int rt_scene_bounds(const rt_scene *s, float bmin[3], float bmax[3]) {
    return api_call([&] {
        if (!s || !bmin || !bmax) throw ApiError(RT_ERR_INVALID, "null argument");
        rt::Vec3 lo, hi;
        if (!s->scene.bounds(lo, hi)) throw ApiError(RT_ERR_INVALID, "scene has no geometry");
        bmin[0] = lo.x; bmin[1] = lo.y; bmin[2] = lo.z;
        bmax[0] = hi.x; bmax[1] = hi.y; bmax[2] = hi.z;
    });
}

int rt_scene_flatten_triangles(const rt_scene *s, rt_triangle *out, uint32_t cap) {
    return api_call([&] {
        if (!s || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        const auto t = s->scene.flatten_triangles();
        if (cap < t.size()) throw ApiError(RT_ERR_INVALID, "capacity below triangle count");
        if (!t.empty()) std::memcpy(out, t.data(), t.size() * sizeof(rt_triangle));
    });
}

int rt_scene_flatten_materials(const rt_scene *s, rt_material *out, uint32_t cap) {
    return api_call([&] {
        if (!s || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        if (cap < s->scene.materials.size()) throw ApiError(RT_ERR_INVALID, "capacity below material count");
        if (!s->scene.materials.empty())
            std::memcpy(out, s->scene.materials.data(), s->scene.materials.size() * sizeof(rt_material));
    });
}

int rt_scene_add_texture(rt_scene *s, uint32_t role, uint32_t width, uint32_t height, const float *rgba,
                         uint32_t *layer) {
    return api_call([&] {
        if (!s || !rgba) throw ApiError(RT_ERR_INVALID, "null argument");
        if (role >= RT_TEX_ROLES) throw ApiError(RT_ERR_INVALID, "texture role out of range");
        if (width == 0 || height == 0 || (uint64_t)width * height > (1ull << 26))
            throw ApiError(RT_ERR_INVALID, "texture size must be non-zero and at most 2^26 texels");
        rt::Texture t;
        t.width = width;
        t.height = height;
        t.rgba.assign(rgba, rgba + (size_t)width * height * 4);
        s->scene.textures[role].push_back(std::move(t));
        if (layer) *layer = (uint32_t)(s->scene.textures[role].size() - 1);
    });
}

int rt_scene_texture_array(const rt_scene *s, uint32_t role, uint8_t *texels, size_t capacity, uint32_t *width,
                           uint32_t *height, uint32_t *layers) {
    return api_call([&] {
        if (!s) throw ApiError(RT_ERR_INVALID, "null scene");
        if (role >= RT_TEX_ROLES) throw ApiError(RT_ERR_INVALID, "texture role out of range");
        const rt::TextureArray a = rt::build_texture_array(s->scene.textures[role]);
        if (width) *width = a.width;
        if (height) *height = a.height;
        if (layers) *layers = a.layers;
        if (!texels) return;
        if (capacity < a.texels.size()) throw ApiError(RT_ERR_INVALID, "capacity below width*height*layers*4");
        std::memcpy(texels, a.texels.data(), a.texels.size());
    });
}

int rt_tonemap_cpu(const float *accum_rgba, size_t n_px, uint8_t *rgba8) {
    return api_call([&] {
        if (n_px && (!accum_rgba || !rgba8)) throw ApiError(RT_ERR_INVALID, "null argument");
        // examples/common/mod.rs:37-47: Reinhard without exposure, clamp (NaN kept),
        // powf(1/2.2) * 255 and a saturating cast to u8 (NaN -> 0); alpha 255
        const float e = 1.0f / 2.2f;
        for (size_t i = 0; i < n_px; i++) {
            for (int k = 0; k < 3; k++) {
                const float v = accum_rgba[4 * i + k];
                float m = v / (1.0f + v);
                if (m == m) m = std::fmin(std::fmax(m, 0.0f), 1.0f);
                const float o = std::pow(m, e) * 255.0f;
                rgba8[4 * i + k] = (o != o || o <= 0.0f) ? 0 : (o >= 255.0f ? 255 : (uint8_t)o);
            }
            rgba8[4 * i + 3] = 255;
        }
    });
}

int rt_srgb_to_linear(float *values, size_t n) {
    return api_call([&] {
        if (n && !values) throw ApiError(RT_ERR_INVALID, "null argument");
        // loader.rs:28-34 in f32; std::pow(float, float) is the C library's powf,
        // which Rust's f32::powf also calls
        for (size_t i = 0; i < n; i++) {
            const float c = values[i];
            values[i] = c <= 0.04045f ? c / 12.92f : std::pow((c + 0.055f) / 1.055f, 2.4f);
        }
    });
}

int rt_scene_apply_descriptor(rt_scene *s, const char *toml_text) {
    return api_call([&] {
        if (!s || !toml_text) throw ApiError(RT_ERR_INVALID, "null argument");
        rt::apply_descriptor(s->scene, toml_text);
    });
}

int rt_scene_light_descs(const rt_scene *s, rt_light_desc *out, uint32_t cap, uint32_t *n) {
    return api_call([&] {
        if (!s) throw ApiError(RT_ERR_INVALID, "null scene");
        const auto &l = s->scene.lights;
        if (n) *n = (uint32_t)l.size();
        if (out) {
            if (cap < l.size()) throw ApiError(RT_ERR_INVALID, "capacity below light count");
            if (!l.empty()) std::memcpy(out, l.data(), l.size() * sizeof(rt_light_desc));
        }
    });
}

int rt_scene_lights_gpu(const rt_scene *s, rt_light *out, uint32_t cap) {
    return api_call([&] {
        if (!s || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        const auto l = s->scene.lights_gpu();
        if (cap < l.size()) throw ApiError(RT_ERR_INVALID, "capacity below light count");
        if (!l.empty()) std::memcpy(out, l.data(), l.size() * sizeof(rt_light));
    });
}

int rt_light_to_gpu(const rt_light_desc *l, rt_light *out) {
    return api_call([&] {
        if (!l || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        *out = rt::light_to_gpu(*l);
    });
}

static void copy_bvh(const rt::BVH &b, rt_bvh_node *nodes, uint32_t ncap, uint32_t *n_nodes, uint32_t *prims,
                     uint32_t pcap, uint32_t *n_prims) {
    if (n_nodes) *n_nodes = (uint32_t)b.nodes.size();
    if (n_prims) *n_prims = (uint32_t)b.prim_indices.size();
    if (nodes) {
        if (ncap < b.nodes.size()) throw ApiError(RT_ERR_INVALID, "node capacity too small");
        if (!b.nodes.empty()) std::memcpy(nodes, b.nodes.data(), b.nodes.size() * sizeof(rt_bvh_node));
    }
    if (prims) {
        if (pcap < b.prim_indices.size()) throw ApiError(RT_ERR_INVALID, "prim capacity too small");
        if (!b.prim_indices.empty()) std::memcpy(prims, b.prim_indices.data(), b.prim_indices.size() * 4);
    }
}

int rt_scene_build_bvh(const rt_scene *s, rt_bvh_node *nodes, uint32_t ncap, uint32_t *n_nodes, uint32_t *prims,
                       uint32_t pcap, uint32_t *n_prims) {
    return api_call([&] {
        if (!s) throw ApiError(RT_ERR_INVALID, "null scene");
        copy_bvh(s->scene.build_bvh(), nodes, ncap, n_nodes, prims, pcap, n_prims);
    });
}

int rt_bvh_build_boxes(const float *centroids, const float *box_min, const float *box_max, uint32_t n,
                       rt_bvh_node *nodes, uint32_t ncap, uint32_t *n_nodes, uint32_t *prims) {
    return api_call([&] {
        if (n && (!centroids || !box_min || !box_max)) throw ApiError(RT_ERR_INVALID, "null argument");
        std::vector<rt::Vec3> c(n);
        std::vector<rt::AABB> b(n);
        for (uint32_t i = 0; i < n; i++) {
            c[i] = rt::Vec3{centroids[3 * i], centroids[3 * i + 1], centroids[3 * i + 2]};
            b[i].min = rt::Vec3{box_min[3 * i], box_min[3 * i + 1], box_min[3 * i + 2]};
            b[i].max = rt::Vec3{box_max[3 * i], box_max[3 * i + 1], box_max[3 * i + 2]};
        }
        copy_bvh(rt::build_bvh(c, b), nodes, ncap, n_nodes, prims, n, nullptr);
    });
}

int rt_camera_default(rt_camera *c) {
    return api_call([&] {
        if (!c) throw ApiError(RT_ERR_INVALID, "null camera");
        *c = rt::camera_default();
    });
}

int rt_camera_to_uniform(const rt_camera *c, uint32_t w, uint32_t h, rt_camera_uniform *out) {
    return api_call([&] {
        if (!c || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        if (w == 0 || h == 0) throw ApiError(RT_ERR_INVALID, "zero resolution");
        *out = rt::camera_to_uniform(*c, w, h);
    });
}

int rt_camera_frame_scene(const rt_scene *s, rt_camera *out) {
    return api_call([&] {
        if (!s || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        *out = rt::camera_frame_scene(s->scene);
    });
}

int rt_camera_orbit(rt_camera *c, float yaw, float pitch) {
    return api_call([&] {
        if (!c) throw ApiError(RT_ERR_INVALID, "null camera");
        rt::camera_orbit(*c, yaw, pitch);
    });
}

int rt_camera_zoom(rt_camera *c, float d) {
    return api_call([&] {
        if (!c) throw ApiError(RT_ERR_INVALID, "null camera");
        rt::camera_zoom(*c, d);
    });
}

int rt_camera_pan(rt_camera *c, float dx, float dy) {
    return api_call([&] {
        if (!c) throw ApiError(RT_ERR_INVALID, "null camera");
        rt::camera_pan(*c, dx, dy);
    });
}

}  // extern "C"
